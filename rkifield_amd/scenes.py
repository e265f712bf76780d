"""Deterministic synthetic workloads (splitmix64 -> PCG32, draw order written down).

* ``spheres_scene``  64 voxelised spheres in the reference layout (one object
  per sphere, per-object dense brick maps, AoS VoxelSample pool).
* ``primitive_field`` 4096 spheres/boxes/capsules with smooth-min groups of 8
  for the sparse two-level scenes.

Sizes follow the workload definitions: ``size`` is the full extent of a
primitive (sphere diameter, box edge scale, capsule length).
"""
import math

import numpy as np

from . import layouts as L
from . import march as M

_M64 = (1 << 64) - 1


def splitmix64(state):
    state = (state + 0x9E3779B97F4A7C15) & _M64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return state, z ^ (z >> 31)


class Pcg32:
    """PCG32 (XSH-RR) seeded from two splitmix64 outputs."""

    def __init__(self, seed):
        s, a = splitmix64(seed & _M64)
        s, b = splitmix64(s)
        self.inc = (b << 1 | 1) & _M64
        self.state = 0
        self.next_u32()
        self.state = (self.state + a) & _M64
        self.next_u32()

    def next_u32(self):
        old = self.state
        self.state = (old * 6364136223846793005 + self.inc) & _M64
        x = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((x >> rot) | (x << ((-rot) & 31))) & 0xFFFFFFFF

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next_u32() >> 8) * (1.0 / 16777216.0))

    def unit_vector(self):
        z = self.uniform(-1.0, 1.0)
        phi = self.uniform(0.0, 2.0 * math.pi)
        r = math.sqrt(max(0.0, 1.0 - z * z))
        return np.array([r * math.cos(phi), r * math.sin(phi), z])

    def quaternion(self):
        """Uniform random rotation (Shoemake)."""
        u1, u2, u3 = self.uniform(), self.uniform(), self.uniform()
        a, b = math.sqrt(1 - u1), math.sqrt(u1)
        return np.array([a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2),
                         b * math.sin(2 * math.pi * u3), b * math.cos(2 * math.pi * u3)])


def quat_to_world_to_local(q):
    """Row-major 3x3 world->local rotation (transpose of the quaternion's matrix)."""
    x, y, z, w = q
    r = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    return r.T.astype(np.float32)


def material_seed(seed):
    return (0x9E3779B97F4A7C15 ^ seed) | 1


class RefScene:
    """Reference-layout scene: pool, packed maps, objects, camera, uniforms."""

    def __init__(self, pool, maps, objects, cam, su, ext, width, height, voxel_size):
        self.pool, self.maps, self.objects = pool, maps, objects
        self.cam, self.su, self.ext = cam, su, ext
        self.width, self.height = width, height
        self.voxel_size = voxel_size

    def aabbs(self):
        return [(o["aabb_min"][:3].copy(), o["aabb_max"][:3].copy()) for o in self.objects]


def spheres_scene(seed=0, n=64, voxel_size=0.05, width=640, height=360, r_range=(0.2, 1.0), half_cube=8.0,
                  cam_dist=24.0, fov=60.0, rel_eps=True):
    """64 spheres: radius U[0.2,1], centre U[-8,8]^3 (draws: r, cx, cy, cz per
    sphere), each voxelised as its own object (margin 4 vs as in
    gen_test_rkf.rs:41), per-voxel material U{0..15}; camera at 24 * u with u
    uniform on the sphere (draws: z, phi), looking at the origin."""
    rng = Pcg32(seed)
    spheres = []
    for _ in range(n):
        r = rng.uniform(*r_range)
        c = [rng.uniform(-half_cube, half_cube) for _ in range(3)]
        spheres.append((np.float32(r), np.array(c, np.float32)))
    cam_dir = rng.unit_vector()
    pools, maps = [], []
    objects = np.zeros(n, L.GPU_OBJECT)
    slot = 0
    off = 0
    vs = np.float32(voxel_size)
    be = np.float32(vs * np.float32(8.0))
    mseed = material_seed(seed)
    for i, (r, c) in enumerate(spheres):
        dims, m, b = M.voxelize_sphere(float(r), float(vs * np.float32(4.0)), float(vs), i, mseed, 0, slot)
        nb = b.size // 512
        o = objects[i]
        inv = np.eye(4, dtype=np.float32)
        inv[3, :3] = -c  # column 3 (translation) of a column-major matrix
        o["inverse_world"] = inv.reshape(-1)
        half = dims.astype(np.float32) * be * np.float32(0.5)
        o["aabb_min"][:3] = c - half
        o["aabb_max"][:3] = c + half
        o["brick_map_offset"] = off
        o["brick_map_dims"] = dims
        o["voxel_size"] = vs
        o["sdf_type"] = L.SDF_TYPE_VOXELIZED
        o["accumulated_scale"] = (1.0, 1.0, 1.0)
        o["object_id"] = i + 1
        # tight local AABB of allocated bricks (grid centred on the origin)
        occ = np.argwhere((m != L.EMPTY_SLOT).reshape(dims[2], dims[1], dims[0]))
        if occ.size:
            lo = occ.min(axis=0)[::-1].astype(np.float32)
            hi = (occ.max(axis=0)[::-1] + 1).astype(np.float32)
            o["geometry_aabb_min"] = lo * be - half
            o["geometry_aabb_max"] = hi * be - half
        pools.append(b)
        maps.append(m)
        slot += nb
        off += m.size
    pool = np.concatenate(pools) if pools else np.zeros(0, L.VOXEL_SAMPLE)
    mp = np.concatenate(maps)
    cam = M.camera_look_at(cam_dir * cam_dist, (0.0, 0.0, 0.0), fov, width, height)
    su = L.scene_uniforms(num_objects=n)
    ext = L.scene_ext(hit_eps_mode=L.EPS_RELATIVE if rel_eps else L.EPS_ABSOLUTE)
    return RefScene(pool, mp, objects, cam, su, ext, width, height, float(vs))


def primitive_field(seed=1, n=4096, half_cube=128.0, size_range=(0.1, 4.0), group=8):
    """4096 primitives: type U{sphere, box, capsule}, size U[0.1,4] (full
    extent), uniform random rotation, centre U[-128,128]^3; draws per
    primitive: type, size, 3 box aspect factors, quaternion (3), centre (3).
    Consecutive groups of 8 are blended with smin."""
    rng = Pcg32(seed)
    prims = np.zeros(n, L.PRIM)
    for i in range(n):
        t = min(int(rng.uniform(0.0, 3.0)), 2)
        s = rng.uniform(*size_range)
        asp = [rng.uniform(0.5, 1.0) for _ in range(3)]
        q = rng.quaternion()
        c = [rng.uniform(-half_cube, half_cube) for _ in range(3)]
        p = prims[i]
        p["type"] = t
        if t == 0:
            p["params"][0] = s * 0.5
        elif t == 1:
            p["params"][:3] = [s * 0.5 * a for a in asp]
        else:
            p["params"][0] = s * 0.25
            p["params"][1] = s * 0.25
        p["rot"] = quat_to_world_to_local(q).reshape(-1)
        p["centre"] = c
        p["group"] = i // group
        p["material"] = i % 16
    return prims


def prim_bounds(prims, pad):
    """World AABB covering every primitive's bounding sphere plus pad."""
    r = np.where(prims["type"] == 0, prims["params"][:, 0],
                 np.where(prims["type"] == 1, np.linalg.norm(prims["params"][:, :3], axis=1),
                          prims["params"][:, 0] + prims["params"][:, 1]))
    lo = (prims["centre"] - r[:, None]).min(axis=0) - pad
    hi = (prims["centre"] + r[:, None]).max(axis=0) + pad
    return lo.astype(np.float32), hi.astype(np.float32)


def mixed_batches(seed=2, n_views=16, n_shadow=16, n_points=32, half_cube=128.0):
    """Config-2 batch mix over a 256 m scene: views with W ~ U{256..1920},
    H ~ U{256..1080} (camera 64 m from the centre on a random direction,
    looking at a point U[-16,16]^3), shadow batches and point batches of
    U{1k..64k} rays/points with origins/points U in the scene box. Draw order:
    all views (W, H, direction, target), then shadow sizes, then point sizes."""
    rng = Pcg32(seed)
    views = []
    for _ in range(n_views):
        w = 256 + int(rng.uniform(0, 1665))
        h = 256 + int(rng.uniform(0, 825))
        d = rng.unit_vector()
        tgt = [rng.uniform(-16, 16) for _ in range(3)]
        views.append((w, h, d * 64.0, tgt))
    shadows = [1024 + int(rng.uniform(0, 64 * 1024 - 1024 + 1)) for _ in range(n_shadow)]
    points = [1024 + int(rng.uniform(0, 64 * 1024 - 1024 + 1)) for _ in range(n_points)]
    nrng = np.random.default_rng(seed)
    shadow_origins = [nrng.uniform(-half_cube, half_cube, (n, 3)).astype(np.float32) for n in shadows]
    point_xyz = [nrng.uniform(-half_cube, half_cube, (n, 3)).astype(np.float32) for n in points]
    return views, shadow_origins, point_xyz


def served_trace(seed=3, n=4000, rate=2000.0, n_scenes=8):
    """Served-stream arrival trace: Poisson arrivals at `rate` requests/s
    (inter-arrival -ln(1-U)/rate); kind by U: < 0.70 a point-query batch of
    U{1k..64k} points, < 0.95 a 512x512 view, else a 1920x1080 view; scene
    U{0..n_scenes-1}; a view's camera sits 64 m from the centre on a random
    direction looking at a point U[-16,16]^3. Draw order per request: gap,
    kind, scene, then the batch size (points) or direction (2) and target (3).
    Returns a list of (arrival_s, kind, scene, size_or_view)."""
    rng = Pcg32(seed)
    t = 0.0
    out = []
    for _ in range(n):
        t += -math.log(1.0 - rng.uniform()) / rate
        u = rng.uniform()
        sc = min(int(rng.uniform(0.0, n_scenes)), n_scenes - 1)
        if u < 0.70:
            out.append((t, "points", sc, 1024 + int(rng.uniform(0, 64 * 1024 - 1024 + 1))))
        else:
            w, h = (512, 512) if u < 0.95 else (1920, 1080)
            d = rng.unit_vector()
            tgt = [rng.uniform(-16, 16) for _ in range(3)]
            out.append((t, "view", sc, (w, h, d * 64.0, tgt)))
    return out

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


def _object_transform(rng, half, scale_range):
    """Translation U[-half, half]^3, uniform rotation, per-axis scale
    U[scale_range] (draws: 3 centre, quaternion (3), 3 scales). Returns
    (centre, R local->world 3x3, scale, inverse_world column-major f32):
    world = T * R * S (flatten_object's order), inverse_world = S^-1 R^T T^-1."""
    c = np.array([rng.uniform(-half, half) for _ in range(3)])
    q = rng.quaternion()
    sc = np.array([rng.uniform(*scale_range) for _ in range(3)])
    rot = quat_to_world_to_local(q).astype(np.float64).T  # local -> world
    w = np.eye(4)
    w[:3, :3] = rot @ np.diag(sc)
    w[:3, 3] = c
    inv = np.linalg.inv(w)
    return c, rot, sc, inv.T.astype(np.float32).reshape(-1)  # column-major: m[c*4 + r]


def _world_aabb(rot, sc, c, lo, hi):
    corners = np.array([[x, y, z] for x in (lo[0], hi[0]) for y in (lo[1], hi[1]) for z in (lo[2], hi[2])])
    wc = (rot @ (np.diag(sc) @ corners.T)).T + c
    return wc.min(axis=0).astype(np.float32), wc.max(axis=0).astype(np.float32)


def objects_scene(seed=5, n_analytic=12, n_voxel=8, voxel_size=0.05, width=160, height=96, half=2.5,
                  cam_dist=9.0, jitter=(0.31, -0.27), rel_eps=True, scale_range=(0.6, 1.4)):
    """Reference-layout scene that exercises every object path of the march:
    analytic objects cycling through the six primitives (sphere, box,
    capsule, torus, cylinder, plane; ray_march.wgsl:182-264), geometry-first
    voxelised primitives (INTERIOR_SLOT bricks, secondary material and blend
    weight per voxel), each object rotated and scaled non-uniformly
    (inverse_world and accumulated_scale), half of the voxelised ones with a
    geometry AABB; a camera with sub-pixel jitter whose prev_vp is the
    view-projection of a displaced previous pose. One object's material id is
    2^24 + 1, which the f32 round trip of evaluate_object turns into 2^24.
    Draw order: per analytic object its transform then its parameters (a
    plane draws a tilt quaternion first), per voxelised object its transform
    and size, then the camera direction (its y made positive) and the
    previous-pose offset."""
    rng = Pcg32(seed)
    n = n_analytic + n_voxel
    objects = np.zeros(n, L.GPU_OBJECT)
    pools, maps = [], []
    slot = off = 0
    vs = np.float32(voxel_size)
    be = np.float32(vs * np.float32(8.0))
    mseed = material_seed(seed)
    for i in range(n_analytic):
        o = objects[i]
        c, rot, sc, inv = _object_transform(rng, half, scale_range)
        t = i % 6
        if t == L.PRIM_SPHERE:
            prm = (rng.uniform(0.3, 0.7), 0.0, 0.0)
            ext = (prm[0],) * 3
        elif t == L.PRIM_BOX:
            prm = tuple(rng.uniform(0.15, 0.6) for _ in range(3))
            ext = prm
        elif t == L.PRIM_CAPSULE:
            prm = (rng.uniform(0.15, 0.35), rng.uniform(0.2, 0.6), 0.0)
            ext = (prm[0], prm[0] + prm[1], prm[0])
        elif t == L.PRIM_TORUS:
            prm = (rng.uniform(0.35, 0.6), rng.uniform(0.08, 0.2), 0.0)
            ext = (prm[0] + prm[1], prm[1], prm[0] + prm[1])
        elif t == L.PRIM_CYLINDER:
            prm = (rng.uniform(0.2, 0.5), rng.uniform(0.2, 0.6), 0.0)
            ext = (prm[0], prm[1], prm[0])
        else:
            # a ground plane below the other objects (camera above it): the
            # object's transform keeps its scale and a small tilt, its
            # normal (params) leans off local +y; its AABB bounds the slab
            # the tile cull sees
            c = np.array([c[0] * 0.3, -half - 0.8, c[2] * 0.3])
            tilt = rng.quaternion()
            tilt = np.array([tilt[0] * 0.15, tilt[1] * 0.15, tilt[2] * 0.15, 1.0])
            tilt /= np.linalg.norm(tilt)
            rot = quat_to_world_to_local(tilt).astype(np.float64).T
            w = np.eye(4)
            w[:3, :3] = rot @ np.diag(sc)
            w[:3, 3] = c
            inv = np.linalg.inv(w).T.astype(np.float32).reshape(-1)
            prm = (rng.uniform(-0.2, 0.2), 1.0, rng.uniform(-0.2, 0.2))
            ext = (3.0, 0.3, 3.0)
        o["inverse_world"] = inv
        lo, hi = _world_aabb(rot, sc, c, -np.array(ext), np.array(ext))
        o["aabb_min"][:3], o["aabb_max"][:3] = lo, hi
        o["sdf_type"] = L.SDF_TYPE_ANALYTICAL
        o["primitive_type"] = t
        o["sdf_params"][:3] = prm
        o["accumulated_scale"] = sc
        o["material_id"] = (1 << 24) + 1 if i == 1 else 1 + (i * 7) % 63
        o["object_id"] = i + 1
    for j in range(n_voxel):
        i = n_analytic + j
        o = objects[i]
        c, rot, sc, inv = _object_transform(rng, half, scale_range)
        pr = np.zeros((), L.PRIM)
        pr["type"] = j % 3
        size = rng.uniform(0.55, 0.9)
        pr["params"][:3] = (size, size * 0.75, size * 0.6) if j % 3 == 1 else (size, size * 0.7, 0.0) \
            if j % 3 == 2 else (size, 0.0, 0.0)
        if j % 3 == 2:
            pr["params"][0] = size * 0.55
        pr["rot"] = np.eye(3, dtype=np.float32).reshape(-1)
        dims, m, b = M.voxelize_prim(pr, float(vs * np.float32(4.0)), float(vs),
                                     L.VOX_GEOMETRY_FIRST | L.VOX_BLEND, i, mseed, 0, slot)
        nb = b.size // 512
        o["inverse_world"] = inv
        hgrid = dims.astype(np.float32) * be * np.float32(0.5)
        lo, hi = _world_aabb(rot, sc, c, -hgrid, hgrid)
        o["aabb_min"][:3], o["aabb_max"][:3] = lo, hi
        o["brick_map_offset"] = off
        o["brick_map_dims"] = dims
        o["voxel_size"] = vs
        o["sdf_type"] = L.SDF_TYPE_VOXELIZED
        o["accumulated_scale"] = sc
        o["object_id"] = i + 1
        occ = np.argwhere((m != L.EMPTY_SLOT).reshape(dims[2], dims[1], dims[0]))
        if j % 2 == 0 and occ.size:
            glo = occ.min(axis=0)[::-1].astype(np.float32)
            ghi = (occ.max(axis=0)[::-1] + 1).astype(np.float32)
            o["geometry_aabb_min"] = glo * be - hgrid
            o["geometry_aabb_max"] = ghi * be - hgrid
        pools.append(b)
        maps.append(m)
        slot += nb
        off += m.size
    cam_dir = rng.unit_vector()
    cam_dir[1] = abs(cam_dir[1]) + 0.35  # above the ground planes
    cam_dir /= np.linalg.norm(cam_dir)
    prev_off = np.array([rng.uniform(-0.4, 0.4) for _ in range(3)])
    pool = np.concatenate(pools) if pools else np.zeros(0, L.VOXEL_SAMPLE)
    mp = np.concatenate(maps) if maps else np.zeros(0, np.uint32)
    eye = cam_dir * cam_dist
    cam = M.camera_look_at(eye, (0.0, 0.0, 0.0), 60.0, width, height)
    prev = M.camera_look_at(eye + prev_off, (0.1, -0.05, 0.0), 60.0, width, height)
    cam["prev_vp"] = prev["prev_vp"]
    cam["jitter"][:2] = jitter
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


def config1_view(width=1920, height=1080, shadows=True, cull=True):
    """The 1080p view of the 4096-primitive scene: camera 64 m from the centre
    on the direction Pcg32(1001).unit_vector(), looking at the origin, fov 60;
    128 near steps, hit eps 1e-3 t, far 100 m; one shadow ray per hit toward
    normalize(0.5, 1, 0.3); per-tile ray intervals. Returns (cam, su, ext)."""
    rng = Pcg32(1001)
    cam = M.camera_look_at(rng.unit_vector() * 64.0, (0.0, 0.0, 0.0), 60.0, width, height)
    su = L.scene_uniforms(num_objects=1, max_steps=128, max_distance=100.0, hit_threshold=1e-3)
    light = np.array([0.5, 1.0, 0.3])
    ext = L.scene_ext(hit_eps_mode=L.EPS_RELATIVE, shadow_enable=1 if shadows else 0,
                      light_dir=light / np.linalg.norm(light), march_flags=L.MARCH_TILE_CULL if cull else 0)
    return cam, su, ext


def terrain_view(prim, width=3840, height=2160, cull=True):
    """The terrain frame's camera: 14 m above the terrain at (0, -60) looking
    at the terrain point (0, 25); limits and shadows as config1_view."""
    eye = (0.0, float(terrain_height(prim, 0.0, -60.0)) + 14.0, -60.0)
    tgt = (0.0, float(terrain_height(prim, 0.0, 25.0)), 25.0)
    cam = M.camera_look_at(eye, tgt, 60.0, width, height)
    _, su, ext = config1_view(cull=cull)
    return cam, su, ext


def terrain_edits(prim, rng, n=32):
    """One frame of CSG edits on the terrain in front of the camera: per edit
    x U[-40, 40], z U[-20, 70], radius U[0.5, 3], centre height terrain +
    U[-0.5, 0.5] (draws in that order); even edits smooth-union, odd ones
    smooth-subtract, blend 0.3 r, smooth falloff, material 1 + k % 15."""
    ops = []
    for k in range(n):
        x, z = rng.uniform(-40.0, 40.0), rng.uniform(-20.0, 70.0)
        r = rng.uniform(0.5, 3.0)
        y = float(terrain_height(prim, x, z)) + rng.uniform(-0.5, 0.5)
        kind = L.EDIT_SMOOTH_UNION if k % 2 == 0 else L.EDIT_SMOOTH_SUBTRACT
        ops.append(L.edit_op((x, y, z), kind, L.SHAPE_SPHERE, (r, 0.0, 0.0), 1.0, 0.3 * r, L.FALLOFF_SMOOTH,
                             1 + k % 15))
    return np.array(ops, L.EDIT)


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


def _mix64_np(z):
    z = z + np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def terrain_height(prim, x, z):
    """Height of a type-3 terrain primitive at world (x, z), vectorised
    (float32 steps in the order the builder uses; used to place rocks,
    cameras and edits)."""
    x = np.asarray(x, np.float32)
    z = np.asarray(z, np.float32)
    u = x - np.float32(prim["centre"][0])
    w = z - np.float32(prim["centre"][2])
    h = np.zeros(np.broadcast(u, w).shape, np.float32)
    amp, freq = np.float32(1.0), np.float32(prim["params"][1])
    with np.errstate(over="ignore"):
        for i in range(6):
            seed = np.uint64((int(prim["seed"]) + i) & 0xFFFFFFFF)
            uu, vv = u * freq, w * freq
            fu, fv = np.floor(uu), np.floor(vv)
            iu, iv = fu.astype(np.int64), fv.astype(np.int64)
            tu, tv = uu - fu, vv - fv
            su = tu * tu * (np.float32(3.0) - np.float32(2.0) * tu)
            sv = tv * tv * (np.float32(3.0) - np.float32(2.0) * tv)

            def lat(a, b):
                key = (a.astype(np.uint64) & np.uint64(0xFFFFFFFF)) | ((b.astype(np.uint64) & np.uint64(0xFFFFFFFF)) << np.uint64(32))
                hh = _mix64_np(key ^ (seed * np.uint64(0x9E3779B97F4A7C15)))
                return (hh >> np.uint64(40)).astype(np.float32) * np.float32(1.0 / 8388608.0) - np.float32(1.0)
            c00, c10, c01, c11 = lat(iu, iv), lat(iu + 1, iv), lat(iu, iv + 1), lat(iu + 1, iv + 1)
            a = c00 + (c10 - c00) * su
            b = c01 + (c11 - c01) * su
            h = h + amp * (a + (b - a) * sv)
            amp = amp * np.float32(0.5)
            freq = freq * np.float32(2.0)
    return np.float32(prim["params"][0]) * h + np.float32(prim["centre"][1])


def terrain_scene(seed=4, half=500.0, n_rocks=20000, amplitude=20.0, freq=1.0 / 160.0, scale=1.0,
                  rock_size=(0.3, 2.5), group=8):
    """Terrain workload: an fBm heightfield (type 3, 6 octaves, amplitude
    `amplitude` m, base frequency `freq` 1/m, distance scale `scale`) over
    [-half, half]^2 in x/z, and `n_rocks` spheres/boxes (size U[rock_size],
    random rotation) sitting on it, blended with smin in groups of 8. Draws
    (Pcg32(seed)) per rock: type, size, 3 aspect factors, quaternion (3),
    x, z; a rock's centre sits half its size above the terrain. Returns
    (prims, bounds_min, bounds_max)."""
    rng = Pcg32(seed)
    prims = np.zeros(1 + n_rocks, L.PRIM)
    t = prims[0]
    t["type"] = 3
    t["params"][:3] = (amplitude, freq, scale)
    t["rot"] = np.eye(3, dtype=np.float32).reshape(-1)
    t["group"] = 0
    t["seed"] = seed * 7919 + 1
    for i in range(n_rocks):
        ty = min(int(rng.uniform(0.0, 2.0)), 1)
        s = rng.uniform(*rock_size)
        asp = [rng.uniform(0.5, 1.0) for _ in range(3)]
        q = rng.quaternion()
        x = rng.uniform(-half, half)
        z = rng.uniform(-half, half)
        p = prims[1 + i]
        p["type"] = ty
        if ty == 0:
            p["params"][0] = s * 0.5
        else:
            p["params"][:3] = [s * 0.5 * a for a in asp]
        p["rot"] = quat_to_world_to_local(q).reshape(-1)
        p["centre"] = (x, 0.0, z)
        p["group"] = 1 + i // group
        p["material"] = 1 + i % 15
    xs, zs = prims["centre"][1:, 0], prims["centre"][1:, 2]
    hs = terrain_height(t, xs, zs)
    sizes = np.where(prims["type"][1:] == 0, prims["params"][1:, 0], prims["params"][1:, 1])
    prims["centre"][1:, 1] = hs + sizes * np.float32(0.5)
    top = amplitude * 1.96875 + rock_size[1] + 2.0
    lo = np.array([-half, -top, -half], np.float32)
    hi = np.array([half, top, half], np.float32)
    return prims, lo, hi

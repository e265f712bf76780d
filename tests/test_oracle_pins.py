"""Pins of the oracle functions the kernels are checked against, by what the
mathematics fixes rather than by retyping the oracle's formulas:

* analytic primitives (ray_march.wgsl:182-264) through evaluate_object
  (:600-634): points built on each surface and offset along its normal read
  the offset; rotated, non-uniformly scaled objects read the local distance
  times min(scale);
* the march on analytic objects: hit depth against ray-sphere / ray-ellipsoid
  intersection within the hit band, normals against the closed-form
  gradient; the gradient transform on a rotated voxelised sphere;
* the overstep bisection (:941-964): a planar crossing found to MIN_STEP / 2^8;
* motion vectors (:1093-1105) for a camera translated along its right axis;
* smin (sdf.rs:48-51): smin(a, a, k) = a - k/4 and its blend band;
* soft_shadow (shade_common_shading.wgsl:9-33) over a voxelised ground
  plane: unoccluded = 1, blocked = 0, a ray parallel to the plane at height h
  gives k h / t_32 with t_32 = 0.005 + 31 h;
* point queries (sdf_collision.rs:38-56) on the ground plane: distance = y,
  inside = y < 0, gradient (0, 1, 0)."""
import math

import numpy as np

from oracle import oracle_py as O
from rkifield_amd import layouts as L
from rkifield_amd import march as M
from rkifield_amd import scenes


def _rot(seed):
    return scenes.quat_to_world_to_local(scenes.Pcg32(seed).quaternion()).astype(np.float64).T  # local -> world


def _analytic_scene(prims, cam_eye=(0.0, 0.0, 6.0), w=64, h=48, eps_abs=1e-3, target=(0.0, 0.0, 0.0)):
    """prims: list of (type, params, centre, R local->world, scale)."""
    objs = np.zeros(len(prims), L.GPU_OBJECT)
    for i, (t, prm, c, R, sc) in enumerate(prims):
        wm = np.eye(4)
        wm[:3, :3] = R @ np.diag(sc)
        wm[:3, 3] = c
        objs[i]["inverse_world"] = np.linalg.inv(wm).T.astype(np.float32).reshape(-1)
        objs[i]["sdf_type"] = L.SDF_TYPE_ANALYTICAL
        objs[i]["primitive_type"] = t
        objs[i]["sdf_params"][:3] = prm
        objs[i]["accumulated_scale"] = sc
        objs[i]["material_id"] = 3 + i
        objs[i]["object_id"] = 1 + i
        objs[i]["aabb_min"][:3] = np.asarray(c) - 3
        objs[i]["aabb_max"][:3] = np.asarray(c) + 3
    cam = M.camera_look_at(cam_eye, target, 60.0, w, h)
    su = L.scene_uniforms(num_objects=len(prims), hit_threshold=eps_abs)
    ext = L.scene_ext(hit_eps_mode=L.EPS_ABSOLUTE, strategy=L.STRATEGY_BRUTE)
    return O.RefInputs(np.zeros(0, L.VOXEL_SAMPLE), np.zeros(0, np.uint32), objs, cam, su, ext, w, h), cam, objs


def _unit(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def test_analytic_primitives_read_the_offset_along_the_normal():
    rng = np.random.default_rng(1)
    n = 200
    th, ph = rng.uniform(0, 2 * np.pi, n), rng.uniform(0, 2 * np.pi, n)
    dl = rng.uniform(0.02, 0.3, n)  # outside offsets, where the SDFs are exact distances
    cases = []
    # sphere r 0.7: (r + d) u
    u = _unit(rng, n)
    cases.append((L.PRIM_SPHERE, (0.7, 0, 0), u * (0.7 + dl[:, None])))
    # box half extents (0.5, 0.3, 0.4): face points pushed out along the face normal
    hb = np.array([0.5, 0.3, 0.4])
    ax = rng.integers(0, 3, n)
    sg = rng.choice([-1.0, 1.0], n)
    p = rng.uniform(-1, 1, (n, 3)) * hb
    p[np.arange(n), ax] = sg * (hb[ax] + dl)
    cases.append((L.PRIM_BOX, tuple(hb), p))
    # box corners pushed out along the corner diagonal: distance = offset
    cor = rng.choice([-1.0, 1.0], (n, 3))
    cases.append((L.PRIM_BOX, tuple(hb), cor * hb + cor / math.sqrt(3) * dl[:, None]))
    # capsule r 0.25, half height 0.4 along y: side and caps
    r, hh = 0.25, 0.4
    y = rng.uniform(-hh, hh, n)
    side = np.stack([(r + dl) * np.cos(th), y, (r + dl) * np.sin(th)], 1)
    up = _unit(rng, n)
    up[:, 1] = np.abs(up[:, 1]) * sg
    caps = np.stack([np.zeros(n), hh * sg, np.zeros(n)], 1) + up * (r + dl)[:, None]
    cases.append((L.PRIM_CAPSULE, (r, hh, 0), np.concatenate([side, caps])))
    # torus major 0.5, minor 0.15
    R0, r0 = 0.5, 0.15
    rho = R0 + (r0 + dl) * np.cos(ph)
    cases.append((L.PRIM_TORUS, (R0, r0, 0), np.stack([rho * np.cos(th), (r0 + dl) * np.sin(ph), rho * np.sin(th)], 1)))
    # cylinder radius 0.35, half height 0.3: side and caps
    rc, hc = 0.35, 0.3
    yc = rng.uniform(-hc, hc, n)
    side = np.stack([(rc + dl) * np.cos(th), yc, (rc + dl) * np.sin(th)], 1)
    rr = rng.uniform(0, rc, n)
    cap = np.stack([rr * np.cos(th), sg * (hc + dl), rr * np.sin(th)], 1)
    cases.append((L.PRIM_CYLINDER, (rc, hc, 0), np.concatenate([side, cap])))
    # plane with normal params (0.2, 1, -0.3) (normalised by the shader)
    nn = np.array([0.2, 1.0, -0.3]) / np.linalg.norm([0.2, 1.0, -0.3])
    t1 = np.cross(nn, [1.0, 0, 0])
    t1 /= np.linalg.norm(t1)
    t2 = np.cross(nn, t1)
    sd = rng.uniform(-0.5, 0.5, n)
    cases.append((L.PRIM_PLANE, (0.2, 1.0, -0.3),
                  rng.uniform(-2, 2, (n, 1)) * t1 + rng.uniform(-2, 2, (n, 1)) * t2 + sd[:, None] * nn))
    offsets = [dl, dl, dl, np.concatenate([dl, dl]), dl, np.concatenate([dl, dl]), sd]
    for (t, prm, pts), want in zip(cases, offsets):
        # identity transform, then rotated + non-uniformly scaled
        for R, sc, c in ((np.eye(3), np.ones(3), np.zeros(3)), (_rot(t + 3), np.array([0.7, 1.3, 1.1]),
                                                                np.array([0.4, -0.2, 1.0]))):
            inp, _, objs = _analytic_scene([(t, prm, c, R, sc)])
            wp = (R @ (np.diag(sc) @ pts.T)).T + c
            d, m = O.ref_eval_object(inp, 0, wp, prec=64)
            inv = objs[0]["inverse_world"].reshape(4, 4).T.astype(np.float64)
            tol = 1e-6 * (1 + np.abs(want))  # inverse_world is stored in f32
            assert np.allclose(d, want * sc.min(), atol=2e-6, rtol=1e-5), (t, np.abs(d - want * sc.min()).max())
            assert np.all(m == 3), t
            assert tol.max() < 1e-5 and np.isfinite(inv).all()
    # inside a sphere the distance is minus the depth; inside a box minus the
    # distance to the nearest face
    inp, _, _ = _analytic_scene([(L.PRIM_SPHERE, (0.7, 0, 0), np.zeros(3), np.eye(3), np.ones(3))])
    d, _ = O.ref_eval_object(inp, 0, u * 0.3, prec=64)
    assert np.allclose(d, -0.4, atol=1e-7)
    inp, _, _ = _analytic_scene([(L.PRIM_BOX, tuple(hb), np.zeros(3), np.eye(3), np.ones(3))])
    q = rng.uniform(-1, 1, (n, 3)) * hb * 0.9
    d, _ = O.ref_eval_object(inp, 0, q, prec=64)
    assert np.allclose(d, -np.min(hb - np.abs(q), axis=1), atol=1e-7)


def _ray_dirs(cam, w, h):
    px, py = np.meshgrid(np.arange(w), np.arange(h))
    ux = (px + 0.5 + cam["jitter"][0]) / w
    uy = (py + 0.5 + cam["jitter"][1]) / h
    nx, ny = ux * 2 - 1, 1 - uy * 2
    d = cam["forward"][:3] + nx[..., None] * cam["right"][:3] + ny[..., None] * cam["up"][:3]
    return (d / np.linalg.norm(d, axis=-1, keepdims=True)).reshape(-1, 3)


def test_march_on_ellipsoids_meets_ray_intersection_and_closed_form_normals():
    """Sphere-trace hits of a rotated, non-uniformly scaled sphere lie within
    the hit band before the exact ray-ellipsoid intersection, and the
    central-difference normals (eps 10 hit_threshold) match the closed-form
    gradient R S^-1 l / |l| to 2e-3 rad; a uniform scale gives the sphere."""
    for sc in (np.array([1.2, 1.2, 1.2]), np.array([0.7, 1.3, 1.0])):
        R, c, r = _rot(9), np.array([0.2, -0.1, 0.3]), 0.8
        inp, cam, objs = _analytic_scene([(L.PRIM_SPHERE, (r, 0, 0), c, R, sc)], cam_eye=(0.3, 0.6, 3.2))
        res = O.ref_trace(inp, prec=64)
        hit = res["hit"] == 1
        assert hit.sum() > 300
        o = cam["position"][:3].astype(np.float64)
        dirs = _ray_dirs(cam, 64, 48)
        inv = objs[0]["inverse_world"].reshape(4, 4).T.astype(np.float64)
        A = inv[:3, :3]
        ol = A @ o + inv[:3, 3]
        checked, worst = 0, 0.0
        for i in np.flatnonzero(hit):
            dl = A @ dirs[i]
            a, b, cc = dl @ dl, 2 * ol @ dl, ol @ ol - r * r
            disc = b * b - 4 * a * cc
            if disc <= 0:
                # a near miss stopped by the hit band: the ray passes within
                # hit_threshold / min(scale) of the local sphere
                tm = -b / (2 * a)
                assert np.linalg.norm(ol + tm * dl) - r < 1e-3 / sc.min() + 1e-6
                continue
            t_star = (-b - math.sqrt(disc)) / (2 * a)
            p = o + dirs[i] * t_star
            lp = A @ p + inv[:3, 3]
            n_exact = R @ (lp / np.linalg.norm(lp) / sc)
            n_exact /= np.linalg.norm(n_exact)
            cosi = abs(n_exact @ dirs[i])
            if cosi < 0.3:
                continue
            # the march stops once (|l| - r) min(scale) < hit_threshold; along
            # the local ray |l| - r falls at rate |dl . l*/|l*|| near t*
            rate = abs(dl @ (lp / np.linalg.norm(lp)))
            band = 1e-3 / sc.min() / rate * 1.01 + 1e-6
            assert -1e-6 <= t_star - res["t"][i] <= band, (i, t_star, res["t"][i])
            # closed-form gradient at the hit point itself
            lh = A @ res["pos"][i] + inv[:3, 3]
            n_hit = R @ (lh / np.linalg.norm(lh) / sc)
            n_hit /= np.linalg.norm(n_hit)
            ang = math.acos(min(1.0, abs(n_hit @ res["normal"][i])))
            worst = max(worst, ang)
            assert ang < 2e-3, (i, ang)
            checked += 1
        assert checked > 200, checked


def test_gradient_transform_on_a_rotated_voxelised_sphere():
    """The normal of a voxelised sphere (voxelize_sdf) under a rotation and a
    uniform scale is transpose(inverse_world) of its local gradient: it points
    along (p - c) / |p - c| (a transposed or untransformed gradient fails)."""
    vs = 0.05
    dims, m, b = M.voxelize_sphere(0.8, vs * 4, vs, 0, 0, 7, 0)
    R, c, s = _rot(21), np.array([0.3, 0.1, -0.2]), 1.3
    wm = np.eye(4)
    wm[:3, :3] = R * s
    wm[:3, 3] = c
    obj = np.zeros(1, L.GPU_OBJECT)
    obj["inverse_world"] = np.linalg.inv(wm).T.astype(np.float32).reshape(-1)
    be = np.float32(vs * 8)
    obj["brick_map_dims"] = dims
    obj["voxel_size"] = vs
    obj["sdf_type"] = L.SDF_TYPE_VOXELIZED
    obj["accumulated_scale"] = (s, s, s)
    obj["object_id"] = 2
    cam = M.camera_look_at((1.0, 2.5, 3.5), c, 60.0, 64, 48)
    inp = O.RefInputs(b, m, obj, cam, L.scene_uniforms(num_objects=1),
                      L.scene_ext(hit_eps_mode=L.EPS_ABSOLUTE, strategy=L.STRATEGY_BRUTE), 64, 48)
    res = O.ref_trace(inp, prec=64)
    hit = res["hit"] == 1
    assert hit.sum() > 300
    radial = res["pos"][hit] - c
    radial /= np.linalg.norm(radial, axis=1, keepdims=True)
    ang = np.arccos(np.clip(np.sum(radial * res["normal"][hit], axis=1), -1, 1))
    assert np.median(ang) < 0.02 and np.percentile(ang, 95) < 0.06
    # depth: the hit lies on the world sphere of radius 0.8 s within O(vs s)
    assert np.abs(np.linalg.norm(res["pos"][hit] - c, axis=1) - 0.8 * s).max() < 0.5 * vs * s


def test_bisection_finds_a_planar_crossing():
    """A plane y = -0.5 (exact SDF) marched with hit_threshold 1e-4 < MIN_STEP:
    rays whose last step overshoots bisect [prev_t, t] (width MIN_STEP) 8
    times, so t lies within MIN_STEP / 2^9 of the exact crossing."""
    inp, cam, _ = _analytic_scene([(L.PRIM_PLANE, (0, 1, 0), np.array([0, -0.5, 0]), np.eye(3), np.ones(3))],
                                  cam_eye=(0.0, 1.5, 2.0), eps_abs=1e-4, target=(0.0, -0.5, -1.0))
    res = O.ref_trace(inp, prec=64)
    dirs = _ray_dirs(cam, 64, 48)
    o = cam["position"][:3].astype(np.float64)
    hit = res["hit"] == 1
    bis = hit & ((res["flags"] & 8) != 0)
    assert bis.sum() > 100 and (hit & ~bis).sum() > 0
    t_star = (-0.5 - o[1]) / dirs[:, 1]
    assert np.abs(res["t"][bis] - t_star[bis]).max() <= 5e-4 / 512 + 1e-7
    # rays that stop without bisecting lie before the plane within the threshold
    nb = hit & ~bis
    gap = (t_star[nb] - res["t"][nb]) * np.abs(dirs[nb, 1])
    assert np.all(gap >= -1e-9) and np.all(gap < 1e-4 + 1e-9)
    assert np.all(res["packed_material"][hit] == (3 | (1 << 24)))


def test_motion_vectors_for_a_translated_camera():
    """prev_vp of the camera moved by delta along its right axis: a hit
    point with view-space coordinates (x, y, z) moves to
    u_prev = 0.5 + 0.5 P00 (x - delta) / -z, v_prev = 0.5 - 0.5 P11 y / -z, and
    motion = pixel-centre uv - prev uv (no jitter in the current uv)."""
    w, h, fov = 64, 48, 60.0
    eye, tgt = np.array([0.3, 1.2, 4.0]), np.array([0.0, 0.0, 0.0])
    inp, cam, _ = _analytic_scene([(L.PRIM_SPHERE, (1.0, 0, 0), np.zeros(3), np.eye(3), np.ones(3))],
                                  cam_eye=eye, w=w, h=h)
    right = cam["right"][:3].astype(np.float64)
    right /= np.linalg.norm(right)
    delta = 0.25
    prev = M.camera_look_at(eye + delta * right, tgt + delta * right, fov, w, h)
    cam["prev_vp"] = prev["prev_vp"]
    cam["jitter"][:2] = (0.2, -0.3)
    inp, cam2, _ = _analytic_scene([(L.PRIM_SPHERE, (1.0, 0, 0), np.zeros(3), np.eye(3), np.ones(3))],
                                   cam_eye=eye, w=w, h=h)
    inp.s.cam = O._bytes(cam, 144)
    res = O.ref_trace(inp, prec=64)
    hit = np.flatnonzero(res["hit"] == 1)
    assert hit.size > 300
    f = cam["forward"][:3].astype(np.float64)
    up = cam["up"][:3].astype(np.float64)
    up /= np.linalg.norm(up)
    P11 = 1.0 / math.tan(math.radians(fov) / 2)
    P00 = P11 / (w / h)
    rel = res["pos"][hit] - eye
    x, y, z = rel @ right, rel @ up, -(rel @ f)
    up_ = 0.5 + 0.5 * P00 * (x - delta) / -z
    vp_ = 0.5 - 0.5 * P11 * y / -z
    px, py = hit % w, hit // w
    mx, my = (px + 0.5) / w - up_, (py + 0.5) / h - vp_
    assert np.abs(res["motion"][hit, 0] - mx).max() < 2e-6
    assert np.abs(res["motion"][hit, 1] - my).max() < 2e-6
    assert np.abs(mx).max() > 0.01


def test_smin_closed_forms():
    for a, k in ((0.3, 0.1), (-0.2, 0.15), (1.0, 0.5)):
        assert math.isclose(O.lib().orc_smin(a, a, k), np.float32(a) - np.float32(k) / 4, rel_tol=0, abs_tol=1e-7)
    # inside the band |a - b| < k: min(a, b) - h^2 k / 4, h = (k - |a - b|) / k
    rng = np.random.default_rng(2)
    for _ in range(50):
        a, b, k = rng.uniform(-1, 1), 0.0, rng.uniform(0.05, 0.5)
        b = a + rng.uniform(-k, k)
        hh = (k - abs(a - b)) / k
        assert math.isclose(O.lib().orc_smin(a, b, k), min(a, b) - hh * hh * k / 4, abs_tol=2e-6)
        # outside the band smin is min
        b2 = a + k * 1.5
        assert O.lib().orc_smin(a, b2, k) == np.float32(min(a, b2))


def _ground_plane(vs=0.05, half=(1.6, 1.2, 1.6), mat=1):
    bws = vs * 8
    dims = [int(np.ceil(2 * hh / bws)) for hh in half]
    origin = np.array([-d * bws / 2 for d in dims], np.float32)
    m = np.full(dims[0] * dims[1] * dims[2], L.EMPTY_SLOT, np.uint32)
    bricks = []
    for bz in range(dims[2]):
        for by in range(dims[1]):
            for bx in range(dims[0]):
                cy = origin[1] + by * bws + bws / 2
                if abs(cy) >= 1.8 * bws:
                    continue
                m[bx + by * dims[0] + bz * dims[0] * dims[1]] = len(bricks)
                v = np.arange(512)
                yv = origin[1] + by * bws + ((v >> 3) & 7) * vs + vs / 2
                b = np.zeros(512, L.VOXEL_SAMPLE)
                b["word0"] = O.f32_to_f16_bits(yv.astype(np.float32)).astype(np.uint32) | (mat << 16)
                bricks.append(b)
    dims_ = [(d + 7) // 8 * 8 for d in dims]
    cells = [(d + 7) // 8 for d in dims]
    l2 = np.full(cells[0] * cells[1] * cells[2], L.EMPTY_SLOT, np.uint32)
    l1 = []
    for bz in range(dims[2]):
        for by in range(dims[1]):
            for bx in range(dims[0]):
                e = m[bx + by * dims[0] + bz * dims[0] * dims[1]]
                if e == L.EMPTY_SLOT:
                    continue
                cc = (bx >> 3) + (by >> 3) * cells[0] + (bz >> 3) * cells[0] * cells[1]
                if l2[cc] == L.EMPTY_SLOT:
                    l2[cc] = len(l1)
                    l1.append(np.full(512, L.EMPTY_SLOT, np.uint32))
                l1[l2[cc]][(bx & 7) + (by & 7) * 8 + (bz & 7) * 64] = e
    del dims_
    b = np.concatenate(bricks)
    return O.SparseInputs(origin, np.float32(vs), cells, l2, np.concatenate(l1), (b["word0"] & 0xFFFF).astype(np.uint16),
                          ((b["word0"] >> 16) & 0x3F).astype(np.uint8))


def test_soft_shadow_closed_forms():
    sp = _ground_plane()
    up = np.array([0.0, 1.0, 0.0], np.float32)
    rng = np.random.default_rng(4)
    xz = rng.uniform(-0.5, 0.5, (64, 2))
    # unoccluded: from above the plane straight up -> exactly 1
    o = np.stack([xz[:, 0], np.full(64, 0.1), xz[:, 1]], 1).astype(np.float32)
    assert np.all(O.sparse_shadow(sp, o, up) == 1.0)
    # toward the plane -> blocked (0)
    down = np.array([0.0, -1.0, 0.0], np.float32)
    assert np.all(O.sparse_shadow(sp, o, down) == 0.0)
    oblique = np.array([0.6, -0.8, 0.0], np.float32)
    assert np.all(O.sparse_shadow(sp, o * [0.2, 1, 0.2], oblique) == 0.0)
    # parallel to the plane at height h (0.005 < h < 0.02): 32 near steps of
    # h from t = 0.005 -> shadow = k h / (0.005 + 31 h)
    par = np.array([1.0, 0.0, 0.0], np.float32)
    for hgt in (0.008, 0.012, 0.018):
        o = np.stack([np.full(16, -0.8), np.full(16, hgt), np.linspace(-0.5, 0.5, 16)], 1).astype(np.float32)
        got = O.sparse_shadow(sp, o, par, k=16.0)
        want = 16.0 * hgt / (0.005 + 31 * hgt)
        assert np.allclose(got, want, rtol=3e-3), (hgt, got, want)
    # k scales the penumbra: shadow(k=4) = shadow(k=16) / 4 below 1
    o = np.stack([np.full(4, -0.8), np.full(4, 0.012), np.zeros(4)], 1).astype(np.float32)
    assert np.allclose(O.sparse_shadow(sp, o, par, k=4.0), O.sparse_shadow(sp, o, par, k=16.0) / 4, rtol=1e-6)


def test_point_query_on_the_ground_plane():
    sp = _ground_plane()
    rng = np.random.default_rng(6)
    q = np.stack([rng.uniform(-1.2, 1.2, 400), rng.uniform(-0.3, 0.3, 400), rng.uniform(-1.2, 1.2, 400)],
                 1).astype(np.float32)
    for prec in (32, 64):
        r = O.sparse_query(sp, q, prec=prec)
        assert np.allclose(r["distance"], q[:, 1], atol=2e-4)
        assert np.array_equal(r["inside"] == 1, r["distance"] < 0)
        assert np.array_equal(r["inside"] == 1, q[:, 1] < 0) or (np.abs(q[:, 1]) < 2e-4).any()
        assert np.allclose(r["gradient"], [0.0, 1.0, 0.0], atol=2e-3)
        assert np.all(r["material"] == 1)


def _cast(o, d, r, max_t=5.0):
    c = np.zeros((), L.SPHERE_CAST)
    c["origin"], c["radius"], c["dir"], c["max_t"] = o, r, np.asarray(d) / np.linalg.norm(d), max_t
    return c


def test_sphere_cast_closed_forms_on_the_ground_plane():
    """Sphere casts (no reference counterpart) over the voxelised plane y = 0:
    a ball of radius r dropped from height h stops with its centre at y = r
    (t = (h - r) / |dir.y| for an oblique cast), normal (0, 1, 0); a cast
    parallel to the plane above r misses; the contact is never past the
    exact one (the distance used is a lower bound)."""
    sp = _ground_plane()
    casts, want = [], []
    for r in (0.05, 0.12, 0.3):
        for h in (0.5, 0.9):
            for d in ((0.0, -1.0, 0.0), (0.5, -1.0, 0.2), (-0.3, -1.0, -0.4)):
                dn = np.asarray(d) / np.linalg.norm(d)
                casts.append(_cast((0.1, h, -0.2), d, r))
                want.append((h - r) / -dn[1])
    casts = np.array(casts, L.SPHERE_CAST)
    for prec in (32, 64):
        res = O.sparse_cast(sp, casts, prec=prec)
        assert np.all(res["hit"] == 1)
        t = res["t"].astype(np.float64)
        assert np.all(t <= np.asarray(want) + 1e-4), (t, want)
        assert np.allclose(t, want, atol=3e-3), np.abs(t - want).max()
        assert np.allclose(res["normal"], [0.0, 1.0, 0.0], atol=2e-3)
    # parallel to the plane at 0.2 above a 0.1 ball: no contact within max_t
    par = np.array([_cast((-1.0, 0.2, 0.0), (1.0, 0.0, 0.0), 0.1, max_t=1.8)], L.SPHERE_CAST)
    r = O.sparse_cast(sp, par)
    assert r["hit"][0] == 0 and r["t"][0] == np.float32(1.8)


def test_sphere_cast_large_radius_is_conservative():
    """A 0.6 m ball (above the 8 vs = 0.4 m cap of the stored field) cast at
    the plane: the contact is reported no later than the exact one and at
    most a narrow band (1.8 bricks) early: the distance bound stops at the
    stored bricks, not at the surface inside them."""
    sp = _ground_plane()
    res = O.sparse_cast(sp, np.array([_cast((0.0, 1.1, 0.0), (0.0, -1.0, 0.0), 0.6)], L.SPHERE_CAST), prec=64)
    assert res["hit"][0] == 1
    exact = 1.1 - 0.6
    assert exact - 1.8 * 0.4 - 1e-3 <= res["t"][0] <= exact + 1e-4
    # from higher up the free space above the band is crossed with large steps
    res = O.sparse_cast(sp, np.array([_cast((0.0, 1.1, 0.0), (0.0, -1.0, 0.0), 0.2)], L.SPHERE_CAST), prec=64)
    assert res["hit"][0] == 1 and abs(res["t"][0] - 0.9) < 3e-3

"""CPU tests of the C ABI library: it loads, exports every symbol the header
declares, and its host-side scene builders agree with the oracle and the
reference's golden fixture. No kernel is launched here."""
import os
import re

import numpy as np

from oracle import oracle_py as O
from rkifield_amd import layouts as L
from rkifield_amd import march as M
from rkifield_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "rkf_march.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(rkfm_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = M.lib()
    names = _declared()
    assert len(names) >= 35
    for n in names:
        assert hasattr(lib, n), n
    assert set(M.exported_symbols()) <= set(names)


def test_layout_sizes():
    assert L.VOXEL_SAMPLE.itemsize == 8
    assert L.GPU_OBJECT.itemsize == 256
    assert L.CAMERA_UNIFORMS.itemsize == 144
    assert L.SCENE_UNIFORMS.itemsize == 16
    assert L.COARSE_FIELD_UNIFORMS.itemsize == 48
    assert L.BVH_NODE.itemsize == 32
    # gpu_object.rs field offsets (table at gpu_object.rs:18-37)
    f = L.GPU_OBJECT.fields
    assert f["aabb_min"][1] == 64 and f["brick_map_offset"][1] == 96 and f["voxel_size"][1] == 112
    assert f["sdf_params"][1] == 132 and f["accumulated_scale"][1] == 148 and f["object_id"][1] == 164
    assert f["geometry_aabb_min"][1] == 172 and f["_padding"][1] == 196
    su = L.scene_uniforms()
    assert su["max_steps"] == 128 and su["max_distance"] == 100.0 and np.isclose(su["hit_threshold"], 0.001)


def test_product_voxelizer_reproduces_golden_fixture():
    data = open(os.path.join(ROOT, "tests", "golden", "test_cross.rkf"), "rb").read()
    h, m, b = O.rkf_load(data)
    dims, m2, b2 = M.voxelize_test_cross()
    assert list(dims) == [8, 4, 8]
    assert np.array_equal(m, m2)
    assert np.array_equal(b.view(np.uint64), b2.view(np.uint64))


def test_product_sphere_voxelizer_matches_oracle_distances():
    import ctypes as C
    for r, vs in ((0.2, 0.05), (1.0, 0.05), (0.73, 0.02)):
        dims, m, b = M.voxelize_sphere(r, vs * 4, vs, 0, 0, 9, 5)
        # the oracle's voxelize_sdf on the same sphere
        lib = O.lib()
        cb = C.CFUNCTYPE(C.c_float, C.POINTER(C.c_float), C.c_void_p, C.POINTER(C.c_uint16))

        def sdf(p, user, mat, r=np.float32(r)):
            mat[0] = 9
            x, y, z = np.float32(p[0]), np.float32(p[1]), np.float32(p[2])
            return float(np.sqrt(np.float32(np.float32(x * x + y * y) + z * z)) - r)
        f = cb(sdf)
        fp = C.cast(f, C.c_void_p).value
        e = np.float32(r) + np.float32(vs) * np.float32(4)
        mn = np.full(3, -e, np.float32)
        mx = np.full(3, e, np.float32)
        d2 = np.zeros(3, np.uint32)
        nb = C.c_uint64(0)
        lib.orc_voxelize_sdf(fp, None, mn.ctypes.data, mx.ctypes.data, C.c_float(vs), C.c_uint32(5),
                             d2.ctypes.data, None, None, C.byref(nb))
        m2 = np.empty(int(np.prod(d2)), np.uint32)
        b2 = np.empty(nb.value * 512, L.VOXEL_SAMPLE)
        lib.orc_voxelize_sdf(fp, None, mn.ctypes.data, mx.ctypes.data, C.c_float(vs), C.c_uint32(5),
                             d2.ctypes.data, m2.ctypes.data, b2.ctypes.data, C.byref(nb))
        assert np.array_equal(dims, d2)
        assert np.array_equal(m, m2)
        assert np.array_equal(b["word0"], b2["word0"])


def _mix64(z):
    z = (z + np.uint64(0x9E3779B97F4A7C15))
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def _oracle_voxelize_prim(prim, margin, vs, geometry, first_slot):
    import ctypes as C
    lib = O.lib()
    pr = np.ascontiguousarray(prim, L.PRIM).reshape(())
    q = pr["params"]
    br = {0: q[0], 1: np.sqrt(np.float32(q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]), 2: q[0] + q[1]}[int(pr["type"])]
    mn = (pr["centre"] - np.float32(br + np.float32(margin))).astype(np.float32)
    mx = (pr["centre"] + np.float32(br + np.float32(margin))).astype(np.float32)
    fn = lib.orc_voxelize_geometry if geometry else lib.orc_voxelize_sdf
    fp = C.cast(lib.orc_prim_sdf, C.c_void_p).value
    d2 = np.zeros(3, np.uint32)
    nb = C.c_uint64(0)
    args = (C.c_float(vs), C.c_uint32(first_slot), d2.ctypes.data)
    fn(C.c_void_p(fp), C.c_void_p(pr.ctypes.data), mn.ctypes.data, mx.ctypes.data, *args, None, None, C.byref(nb))
    m2 = np.empty(int(np.prod(d2)), np.uint32)
    b2 = np.empty(nb.value * 512, L.VOXEL_SAMPLE)
    fn(C.c_void_p(fp), C.c_void_p(pr.ctypes.data), mn.ctypes.data, mx.ctypes.data, *args, m2.ctypes.data,
       b2.ctypes.data, C.byref(nb))
    return d2, m2, b2


def _prims():
    rng = scenes.Pcg32(17)
    out = []
    for t, par in ((0, (0.9, 0, 0)), (1, (0.9, 0.5, 0.7)), (2, (0.45, 0.6, 0))):
        p = np.zeros((), L.PRIM)
        p["type"] = t
        p["params"][:3] = par
        p["rot"] = scenes.quat_to_world_to_local(rng.quaternion()).reshape(-1)
        p["material"] = 5
        out.append(p)
    return out


def test_product_prim_voxelizer_matches_oracle():
    """rkfm_voxelize_prim (voxelize_sdf and the geometry-first classification)
    equals the oracle's restatements word for word, and the classification
    meets its definition, checked independently in numpy on a sphere."""
    for p in _prims():
        for geometry in (False, True):
            flags = L.VOX_GEOMETRY_FIRST if geometry else 0
            dims, m, b = M.voxelize_prim(p, 0.2, 0.05, flags, 0, 0, 5, 3)
            d2, m2, b2 = _oracle_voxelize_prim(p, 0.2, 0.05, geometry, 3)
            assert np.array_equal(dims, d2) and np.array_equal(m, m2)
            assert np.array_equal(b.view(np.uint64), b2.view(np.uint64))
            if geometry:
                assert (m == L.INTERIOR_SLOT).any() and (m < L.INTERIOR_SLOT).any()
    # classification of a sphere of radius 0.9 at vs 0.05 from float32 numpy
    vs, r = np.float32(0.05), np.float32(0.9)
    sph = _prims()[0]
    dims, m, _ = M.voxelize_prim(sph, 0.2, 0.05, L.VOX_GEOMETRY_FIRST)
    bws = vs * np.float32(8)
    go = -(dims.astype(np.float32) * bws * np.float32(0.5))
    band = bws * np.float32(1.8)
    half = vs * np.float32(0.5)
    v = np.arange(512)
    k = 0
    for bz in range(dims[2]):
        for by in range(dims[1]):
            for bx in range(dims[0]):
                bmin = go + np.float32([bx, by, bz]) * bws
                c = bmin + bws * np.float32(0.5)
                dc = np.sqrt(np.float32(c[0] * c[0] + c[1] * c[1]) + c[2] * c[2]) - r
                p = [bmin[a] + (np.float32(vv) * vs + half) for a, vv in enumerate((v & 7, (v >> 3) & 7, v >> 6))]
                dv = np.sqrt((p[0] * p[0] + p[1] * p[1]) + p[2] * p[2]) - r
                if dc < -band or (abs(dc) < band and (dv <= 0).all()):
                    want = "interior"
                elif not abs(dc) < band or (dv > 0).all():
                    want = "empty"
                else:
                    want = "stored"
                e = m[bx + by * dims[0] + bz * dims[0] * dims[1]]
                got = "interior" if e == L.INTERIOR_SLOT else "empty" if e == L.EMPTY_SLOT else "stored"
                assert got == want, (bx, by, bz)
                if got == "stored":
                    assert e == k
                    k += 1
    assert k > 0


def test_product_prim_voxelizer_blend_words():
    """VOX_BLEND: the distance and material bits are the plain ones, the
    secondary (bits 22-27) and blend (word1 byte 3) fields are the second
    hash of the voxel key (restated in numpy), and both cover their range."""
    p = _prims()[1]
    seed = scenes.material_seed(9)
    dims, m, b = M.voxelize_prim(p, 0.2, 0.05, L.VOX_GEOMETRY_FIRST | L.VOX_BLEND, 4, seed, 0, 0)
    _, m0, b0 = M.voxelize_prim(p, 0.2, 0.05, L.VOX_GEOMETRY_FIRST, 4, seed, 0, 0)
    assert np.array_equal(m, m0)
    assert np.array_equal(b["word0"] & 0x3FFFFF, b0["word0"])
    assert np.all(b0["word1"] == 0) and np.all(b0["word0"] >> 22 == 0)
    slots = np.flatnonzero(m < L.INTERIOR_SLOT)
    bz, by, bx = slots // (dims[0] * dims[1]), (slots // dims[0]) % dims[1], slots % dims[0]
    v = np.arange(512)
    gx = (bx[:, None] * 8 + (v & 7)).astype(np.uint64).ravel()
    gy = (by[:, None] * 8 + ((v >> 3) & 7)).astype(np.uint64).ravel()
    gz = (bz[:, None] * 8 + (v >> 6)).astype(np.uint64).ravel()
    key = gx | (gy << np.uint64(21)) | (gz << np.uint64(42))
    with np.errstate(over="ignore"):
        inner = _mix64(key ^ (np.uint64(4) * np.uint64(0xD6E8FEB86659FD93)))
        mat = _mix64(np.uint64(seed) ^ inner) & np.uint64(15)
        h = _mix64(np.uint64(seed) ^ inner ^ np.uint64(0xA24BAED4963EE407))
    order = np.argsort(m[slots])
    w0 = b["word0"].reshape(-1, 512)[m[slots][order]].ravel()
    w1 = b["word1"].reshape(-1, 512)[m[slots][order]].ravel()
    mat, h = mat.reshape(-1, 512)[order].ravel(), h.reshape(-1, 512)[order].ravel()
    assert np.array_equal((w0 >> 16) & 0x3F, mat.astype(np.uint32))
    assert np.array_equal((w0 >> 22) & 0x3F, ((h >> np.uint64(8)) & np.uint64(63)).astype(np.uint32))
    assert np.array_equal(w1 >> 24, ((h >> np.uint64(16)) & np.uint64(255)).astype(np.uint32))
    assert np.unique((w0 >> 22) & 0x3F).size == 64 and np.unique(w1 >> 24).size == 256


def test_per_voxel_material_is_uniform_0_15():
    dims, m, b = M.voxelize_sphere(1.0, 0.2, 0.05, 3, scenes.material_seed(0), 0, 0)
    mat = (b["word0"] >> 16) & 0x3F
    assert mat.max() <= 15
    hist = np.bincount(mat, minlength=16)
    assert hist.min() > 0.8 * hist.mean()


def test_camera_look_at_matches_reference_conventions():
    cam = M.camera_look_at((0.0, 0.0, 5.0), (0, 0, 0), 60.0, 640, 360)
    assert np.allclose(cam["forward"][:3], [0, 0, -1], atol=1e-6)
    hft = np.tan(np.radians(30.0))
    assert np.allclose(cam["right"][:3], [hft * 640 / 360, 0, 0], atol=1e-6)
    assert np.allclose(cam["up"][:3], [0, hft, 0], atol=1e-6)
    assert np.allclose(cam["resolution"], [640, 360]) and np.all(cam["jitter"] == 0)
    # prev_vp projects the look-at target to the screen centre
    vp = cam["prev_vp"].reshape(4, 4).T.astype(np.float64)
    clip = vp @ np.array([0, 0, 0, 1.0])
    assert abs(clip[0] / clip[3]) < 1e-6 and abs(clip[1] / clip[3]) < 1e-6


def test_bvh_covers_every_object_once():
    sc = scenes.spheres_scene(seed=2, n=17, width=16, height=16)
    nodes = M.build_bvh(sc.objects)
    assert nodes.size == 2 * 17 - 1
    leaves = sorted(int(n["right_or_object"]) for n in nodes if n["left"] == L.BVH_INVALID)
    assert leaves == list(range(17))
    for n in nodes:
        if n["left"] != L.BVH_INVALID:
            for ch in (nodes[n["left"]], nodes[n["right_or_object"]]):
                assert np.all(ch["aabb_min"] >= n["aabb_min"]) and np.all(ch["aabb_max"] <= n["aabb_max"])


def test_rng_streams_are_fixed():
    r = scenes.Pcg32(0)
    first = [r.next_u32() for _ in range(4)]
    r2 = scenes.Pcg32(0)
    assert first == [r2.next_u32() for _ in range(4)]
    assert scenes.Pcg32(1).next_u32() != first[0]
    p = scenes.primitive_field(seed=1, n=64)
    assert set(np.unique(p["type"])) <= {0, 1, 2} and p["group"].max() == 7
    assert np.all(np.abs(p["centre"]) <= 128)


def test_tile_cull_oracle_projection_closed_form():
    """A box straight ahead covers the centre tiles only; a box behind the
    camera covers none; a box straddling the camera plane covers all."""
    cam = M.camera_look_at((0, 0, 10), (0, 0, 0), 60.0, 64, 64)
    objs = np.zeros(3, L.GPU_OBJECT)
    objs["sdf_type"] = L.SDF_TYPE_VOXELIZED
    objs[0]["aabb_min"][:3] = [-0.1, -0.1, -0.1]
    objs[0]["aabb_max"][:3] = [0.1, 0.1, 0.1]
    objs[1]["aabb_min"][:3] = [-1, -1, 11]
    objs[1]["aabb_max"][:3] = [1, 1, 12]
    objs[2]["aabb_min"][:3] = [-1, -1, 9]
    objs[2]["aabb_max"][:3] = [1, 1, 11]
    su = L.scene_uniforms(num_objects=3)
    idx, counts = O.tile_cull(cam, su, objs, 64, 64)
    lists = [list(idx[t * 32:t * 32 + counts[t]]) for t in range(16)]
    centre = {5, 6, 9, 10}
    for t, l in enumerate(lists):
        assert (0 in l) == (t in centre)
        assert 1 not in l and 2 in l


def test_coarse_oracle_closed_form():
    objs = np.zeros(1, L.GPU_OBJECT)
    objs[0]["aabb_min"][:3] = [0, 0, 0]
    objs[0]["aabb_max"][:3] = [1, 1, 1]
    cells = O.coarse_populate([-2, 0, 0], [4, 1, 1], 0.5, objs)
    d = O.f16_bits_to_f32(cells)
    # cell centres at x = -1.75, -1.25, -0.75, -0.25 (y, z = 0.25 inside)
    assert np.allclose(d, [1.75, 1.25, 0.75, 0.25], atol=1e-3)

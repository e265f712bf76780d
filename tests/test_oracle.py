"""Pins for the CPU oracle (oracle/) — the reference's own vectors, the golden
fixture, and closed forms. CPU only.

Reference vectors used (file:line in the reference):
  voxel.rs:120-277, brick.rs:189-338, brick_map.rs:322-333, brick_pool.rs:198-540,
  sampling.rs:88-292, sdf.rs:58-167, scenes/test_cross.rkf + examples/gen_test_rkf.rs.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

from oracle import oracle_py as O
from rkifield_amd import layouts as L

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ---------------- binary16 ----------------

def test_f16_matches_ieee_binary16_exhaustively_on_a_sweep():
    rng = np.random.default_rng(0)
    bits = np.concatenate([
        rng.integers(0, 2**32, 2_000_000, dtype=np.uint64).astype(np.uint32),
        np.arange(0x33000000, 0x33800000, 97, dtype=np.uint32),  # subnormal boundary
        np.arange(0x477FE000, 0x47800100, 1, dtype=np.uint32),   # overflow boundary
        np.array([0, 0x80000000, 0x7F800000, 0xFF800000, 0x3F800000, 0x38800000, 0x33800000, 0x33000000],
                 np.uint32),
    ])
    f = bits.view(np.float32)
    finite = np.isfinite(f)
    ours = O.f32_to_f16_bits(f)
    with np.errstate(over="ignore"):  # values past the binary16 range round to infinity, as intended
        ref = f.astype(np.float16).view(np.uint16)
    assert np.array_equal(ours[finite], ref[finite])
    nan = np.isnan(f)
    assert np.all(np.isnan(O.f16_bits_to_f32(ours[nan])))


def test_f16_decode_all_values():
    h = np.arange(65536, dtype=np.uint32).astype(np.uint16)
    ours = O.f16_bits_to_f32(h)
    ref = h.view(np.float16).astype(np.float32)
    both_nan = np.isnan(ours) & np.isnan(ref)
    assert np.array_equal(ours[~both_nan].view(np.uint32), ref[~both_nan].view(np.uint32))


# ---------------- VoxelSample (voxel.rs tests) ----------------

def test_voxel_roundtrip_vectors():
    w0, w1 = O.voxel_new(1.5, 42, 0)
    assert O.f16_bits_to_f32([w0 & 0xFFFF])[0] == 1.5 and (w0 >> 16) & 0x3F == 42 and w1 == 0
    w0, _ = O.voxel_new(-0.5, 1, 0)
    assert O.f16_bits_to_f32([w0 & 0xFFFF])[0] == -0.5
    w0, w1 = O.voxel_new(1.0, 5, 200)
    assert (w0 >> 16) & 0x3F == 5 and w1 >> 24 == 200
    w0, _ = O.voxel_new(0.0, 0xFF, 0)
    assert (w0 >> 16) & 0x3F == 63 and (w0 >> 22) & 0x3F == 0 and w0 & 0xFFFF == 0
    w0, _ = O.voxel_new(0.0, 0x7F, 0)
    assert (w0 >> 16) & 0x3F == 63 and (w0 >> 22) & 0x3F == 0
    w0, _ = O.voxel_new(65504.0, 0, 0)
    assert w0 & 0xFFFF == 0x7BFF  # f16::MAX
    w0, w1 = O.voxel_new(float("inf"), 0, 0)  # VoxelSample::default
    assert w0 == 0x7C00 and w1 == 0
    assert (w0 >> 28) == 0


# ---------------- index math ----------------

def test_brick_index_corners():
    assert O.brick_index(0, 0, 0) == 0
    assert O.brick_index(7, 0, 0) == 7
    assert O.brick_index(0, 7, 0) == 56
    assert O.brick_index(0, 0, 7) == 448
    assert O.brick_index(7, 7, 7) == 511
    assert O.brick_index(8, 0, 0) == -1


def test_map_flatten_vectors():
    d = (3, 4, 5)
    assert O.map_flatten(d, 0, 0, 0) == 0
    assert O.map_flatten(d, 2, 0, 0) == 2
    assert O.map_flatten(d, 0, 1, 0) == 3
    assert O.map_flatten(d, 0, 0, 1) == 12
    assert O.map_flatten(d, 2, 3, 4) == 59
    assert O.map_flatten(d, 3, 0, 0) == -1


def test_pool_allocates_low_slots_first():
    assert O.pool_allocate_sequence(8, 4) == [0, 1, 2, 3]
    assert O.pool_allocate_sequence(2, 3) == [0, 1, -1]


# ---------------- CPU sampling (sampling.rs tests) ----------------

def _brick(fn):
    b = np.zeros(512, L.VOXEL_SAMPLE)
    for z in range(8):
        for y in range(8):
            for x in range(8):
                d, m = fn(x, y, z)
                w0, w1 = O.voxel_new(d, m, 0)
                b[x + 8 * y + 64 * z] = (w0, w1)
    return b


def _centre(x, y, z):
    return [(x + 0.5) / 8, (y + 0.5) / 8, (z + 0.5) / 8]


def test_trilinear_at_voxel_centres():
    b = _brick(lambda x, y, z: (x + y * 0.1 + z * 0.01, (x + 8 * y + 64 * z) % 256))
    for z in range(8):
        for y in range(8):
            for x in range(8):
                stored = O.f16_bits_to_f32([b[x + 8 * y + 64 * z]["word0"] & 0xFFFF])[0]
                assert abs(O.sample_brick_trilinear(b, _centre(x, y, z)) - stored) < 0.02
                assert O.sample_brick_nearest_material(b, _centre(x, y, z)) == ((x + 8 * y + 64 * z) % 256) & 0x3F


def test_trilinear_uniform_midpoint_monotone_clamp():
    b = _brick(lambda x, y, z: (3.5, 1))
    for p in ([0, 0, 0], [0.5, 0.5, 0.5], [1, 1, 1], [0.25, 0.75, 0.1]):
        assert abs(O.sample_brick_trilinear(b, p) - 3.5) < 0.1
    b = _brick(lambda x, y, z: (0.0 if x < 4 else 2.0, 1))
    assert abs(O.sample_brick_trilinear(b, [(3.5 + 0.5) / 8, 0.5, 0.5]) - 1.0) < 0.15
    b = _brick(lambda x, y, z: (float(x), 1))
    prev = O.sample_brick_trilinear(b, [0.0, 0.5, 0.5])
    for i in range(1, 21):
        cur = O.sample_brick_trilinear(b, [i / 20, 0.5, 0.5])
        assert cur >= prev - 0.01
        prev = cur
    g = _brick(lambda x, y, z: (x + y * 0.1 + z * 0.01, 0))
    for p in ([-0.5, 0.5, 0.5], [1.5, 0.5, 0.5], [0.5, -0.1, 1.5]):
        assert np.isfinite(O.sample_brick_trilinear(g, p))


def test_nearest_material_snap_and_clamp():
    b = _brick(lambda x, y, z: (0.0, 10 if (x, y, z) == (3, 3, 3) else 20))
    assert O.sample_brick_nearest_material(b, _centre(3, 3, 3)) == 10
    assert O.sample_brick_nearest_material(b, [3.6 / 8, 3.5 / 8, 3.5 / 8]) == 10
    b = _brick(lambda x, y, z: (0.0, 42))
    assert O.sample_brick_nearest_material(b, [-1, -1, -1]) == 42
    assert O.sample_brick_nearest_material(b, [2, 2, 2]) == 42


# ---------------- analytic SDFs (sdf.rs tests) ----------------

def test_sdf_closed_forms():
    import ctypes as C
    lib = O.lib()
    f3 = C.c_float * 3
    lib.orc_sphere_sdf.restype = C.c_float
    lib.orc_box_sdf.restype = C.c_float
    lib.orc_capsule_sdf.restype = C.c_float
    lib.orc_capsule_sdf.argtypes = [C.c_void_p, C.c_void_p, C.c_float, C.c_void_p]
    lib.orc_sphere_sdf.argtypes = [C.c_void_p, C.c_float, C.c_void_p]
    sph = lambda c, r, p: lib.orc_sphere_sdf(f3(*c), r, f3(*p))
    box = lambda h, p: lib.orc_box_sdf(f3(*h), f3(*p))
    cap = lambda a, b, r, p: lib.orc_capsule_sdf(f3(*a), f3(*b), r, f3(*p))
    assert abs(sph([0, 0, 0], 1, [0, 0, 0]) + 1) < 1e-6
    assert abs(sph([0, 0, 0], 1, [1, 0, 0])) < 1e-6
    assert abs(sph([0, 0, 0], 1, [2, 0, 0]) - 1) < 1e-6
    assert abs(sph([1, 2, 3], 0.5, [1, 2, 3]) + 0.5) < 1e-6
    assert abs(cap([0, 0, 0], [0, 1, 0], 0.5, [0, 0.5, 0]) + 0.5) < 1e-6
    assert abs(cap([0, 0, 0], [0, 1, 0], 0.5, [1.5, 0.5, 0]) - 1.0) < 1e-6
    assert abs(box([1, 1, 1], [0, 0, 0]) + 1) < 1e-6
    assert abs(box([1, 1, 1], [2, 2, 2]) - math.sqrt(3)) < 1e-5
    assert abs(box([1, 1, 1], [0.5, 0, 0]) + 0.5) < 1e-6
    assert abs(O.lib().orc_smin(1.0, 5.0, 0.1) - 1.0) < 1e-7  # |a-b| >= k -> min


# ---------------- golden fixture ----------------

def test_rkf_fixture_header():
    data = open(os.path.join(GOLDEN, "test_cross.rkf"), "rb").read()
    h, m, b = O.rkf_load(data)
    assert h["version"] == 2 and h["lod_count"] == 1 and h["material_ids"] == [6]
    assert list(h["dims"]) == [8, 4, 8] and h["brick_count"] == 224
    assert h["uncompressed_size"] == 256 * 4 + 224 * 4096
    assert np.allclose(h["aabb_max"], [0.96, 0.48, 0.96])


def test_voxelize_sdf_reproduces_test_cross_bit_exactly():
    data = open(os.path.join(GOLDEN, "test_cross.rkf"), "rb").read()
    h, m, b = O.rkf_load(data)
    dims, m2, b2 = O.voxelize_test_cross()
    assert list(dims) == list(h["dims"])
    assert np.array_equal(m, m2)
    assert np.array_equal(b["word0"], b2["word0"])
    assert np.array_equal(b["word1"], b2["word1"])
    # materials: 6 everywhere; slots 0..223 in z-y-x order
    assert np.all(((b["word0"] >> 16) & 0x3F) == 6)
    occ = m[m != L.EMPTY_SLOT]
    assert np.array_equal(occ, np.arange(224))


def test_lz4_rejects_corruption():
    data = bytearray(open(os.path.join(GOLDEN, "test_cross.rkf"), "rb").read())
    data[0:4] = b"XXXX"
    with pytest.raises(ValueError):
        O.rkf_load(bytes(data))


# ---------------- march: closed forms and invariants ----------------

def _single_sphere_scene(r=1.0, vs=0.05, cam=(0.0, 0.0, 3.0), w=48, h=32, rel=False):
    from rkifield_amd import march as M
    dims, m, b = M.voxelize_sphere(r, vs * 4, vs, 0, 0, 7, 0)
    obj = np.zeros(1, L.GPU_OBJECT)
    obj["inverse_world"] = np.eye(4, dtype=np.float32).reshape(-1)
    be = np.float32(vs * 8)
    half = dims.astype(np.float32) * be * np.float32(0.5)
    obj["aabb_min"][0, :3] = -half
    obj["aabb_max"][0, :3] = half
    obj["brick_map_dims"] = dims
    obj["voxel_size"] = vs
    obj["sdf_type"] = L.SDF_TYPE_VOXELIZED
    obj["accumulated_scale"] = (1, 1, 1)
    obj["object_id"] = 3
    cu = M.camera_look_at(cam, (0, 0, 0), 60.0, w, h)
    su = L.scene_uniforms(num_objects=1)
    ext = L.scene_ext(hit_eps_mode=L.EPS_RELATIVE if rel else L.EPS_ABSOLUTE, strategy=L.STRATEGY_BRUTE)
    return O.RefInputs(b, m, obj, cu, su, ext, w, h), cu


def test_voxelized_sphere_depth_matches_ray_sphere_intersection():
    inp, cam = _single_sphere_scene()
    res = O.ref_trace(inp, prec=64)
    o = np.array(cam["position"][:3], np.float64)
    hits = res[res["hit"] == 1]
    assert hits.size > 100
    for r in hits[::7]:
        d = (r["pos"] - o) / np.linalg.norm(r["pos"] - o)
        bq = np.dot(o, d)
        disc = bq * bq - (np.dot(o, o) - 1.0)
        assert disc > -1e-3
        t_exact = -bq - math.sqrt(max(disc, 0.0))
        assert abs(r["t"] - t_exact) < 0.05 * 0.5 + 2e-3  # O(vs) voxelisation error
        n = r["pos"] / np.linalg.norm(r["pos"])
        assert np.dot(n, r["normal"]) > 0.97  # normal close to (p - c)/r
        assert r["material_id"] == 7 and (r["packed_material"] >> 24) == 3


def test_empty_map_misses_everything():
    inp, _ = _single_sphere_scene()
    inp.keep[1][:] = L.EMPTY_SLOT
    res = O.ref_trace(inp, prec=32)
    assert res["hit"].sum() == 0


def test_fp32_twin_agrees_with_fp64_oracle():
    inp, _ = _single_sphere_scene(rel=True)
    a = O.ref_trace(inp, prec=32)
    b = O.ref_trace(inp, prec=64)
    same = (a["hit"] == b["hit"]) & (a["steps"] == b["steps"]) & (a["path"] == b["path"])
    assert same.mean() > 0.98
    hit = same & (a["hit"] == 1)
    rel = np.abs(a["t"][hit] - b["t"][hit]) / b["t"][hit]
    assert rel.max() < 1e-4
    cosang = np.clip(np.sum(a["normal"][hit] * b["normal"][hit], axis=1), -1, 1)
    assert np.arccos(cosang).max() < 1e-3


def test_planted_mistake_in_oracle_input_is_caught():
    """Flip one voxel's sign in the twin's input: the hit/depth output must
    change for the rays that cross it (the checks above can fail)."""
    inp, _ = _single_sphere_scene()
    base = O.ref_trace(inp, prec=32)
    pool = inp.keep[0]
    near = np.where(np.abs(O.f16_bits_to_f32(pool["word0"] & 0xFFFF)) < 0.02)[0]
    pool["word0"][near] ^= 0x8000  # negate near-surface distances
    moved = O.ref_trace(inp, prec=32)
    assert np.any(moved["t"] != base["t"])


# ---------------- sparse path equals the reference field ----------------

def _sparse_from_dense_host(dims, m, b, vs, origin):
    dims = [int(x) for x in dims]
    cells = [(d + 7) // 8 for d in dims]
    l2 = np.full(cells[0] * cells[1] * cells[2], L.EMPTY_SLOT, np.uint32)
    l1 = []
    for bz in range(dims[2]):
        for by in range(dims[1]):
            for bx in range(dims[0]):
                e = m[bx + by * dims[0] + bz * dims[0] * dims[1]]
                if e == L.EMPTY_SLOT:
                    continue
                c = (bx >> 3) + (by >> 3) * cells[0] + (bz >> 3) * cells[0] * cells[1]
                if l2[c] == L.EMPTY_SLOT:
                    l2[c] = len(l1)
                    l1.append(np.full(512, L.EMPTY_SLOT, np.uint32))
                l1[l2[c]][(bx & 7) + (by & 7) * 8 + (bz & 7) * 64] = e
    l1 = np.concatenate(l1) if l1 else np.zeros(0, np.uint32)
    dist = (b["word0"] & 0xFFFF).astype(np.uint16)
    mat = ((b["word0"] >> 16) & 0x3F).astype(np.uint8)
    return O.SparseInputs(origin, vs, cells, l2, l1, dist, mat)


def test_sparse_field_equals_reference_field_at_origin():
    """test_cross.rkf placed at the origin: the sparse field p - origin with
    origin = -grid/2 is bit-identical to sample_voxelized's local + grid/2."""
    data = open(os.path.join(GOLDEN, "test_cross.rkf"), "rb").read()
    h, m, b = O.rkf_load(data)
    vs = np.float32(h["voxel_size"])
    be = vs * np.float32(8.0)
    half = h["dims"].astype(np.float32) * be * np.float32(0.5)
    sp = _sparse_from_dense_host(h["dims"], m, b, vs, -half)
    obj = np.zeros(1, L.GPU_OBJECT)
    obj["inverse_world"] = np.eye(4, dtype=np.float32).reshape(-1)
    obj["brick_map_dims"] = h["dims"]
    obj["voxel_size"] = vs
    obj["sdf_type"] = L.SDF_TYPE_VOXELIZED
    obj["accumulated_scale"] = (1, 1, 1)
    su = L.scene_uniforms(num_objects=1)
    # the full trace of a small view through both layouts must agree bit for bit
    from rkifield_amd import march as M
    cu = M.camera_look_at((0.3, 1.0, 1.6), (0, 0, 0), 60.0, 40, 30)
    inp = O.RefInputs(b, m, obj, cu, su, L.scene_ext(strategy=L.STRATEGY_BRUTE), 40, 30)
    ref = O.ref_trace(inp, prec=32)
    ext = L.scene_ext()
    spr = O.sparse_trace(sp, cu, su, ext, 40, 30, prec=32)
    assert ref["hit"].sum() > 100
    assert np.array_equal(ref["hit"], spr["hit"])
    hit = ref["hit"] == 1
    # DDA changes only free-space steps: depths agree within the hit threshold
    assert np.abs(ref["t"][hit] - spr["t"][hit]).max() < 2e-3
    assert np.array_equal(ref["material_id"][hit], spr["material_id"][hit])


def test_sparse_dda_trace_stays_in_the_hit_band_of_the_plain_trace():
    """The sparse march (free-space DDA) against the reference's plain trace
    (ray_march_brute semantics, :776-858) of the same field, both FP64, over
    test_cross.rkf from four directions with absolute and relative hit
    thresholds: the DDA moves only the steps taken through free space, so
    hit masks and materials are equal, and the hit points differ along the
    surface normal by less than 2 hit thresholds (measured up to 1.57); the
    depth then differs by up to 1.6e-3 relative and up to 2.5 % of the hits
    land in another voxel of the band (this bound, not bit equality, is what
    the sparse path promises against the plain trace)."""
    from rkifield_amd import march as M
    data = open(os.path.join(GOLDEN, "test_cross.rkf"), "rb").read()
    h, m, b = O.rkf_load(data)
    vs = np.float32(h["voxel_size"])
    half = h["dims"].astype(np.float32) * (vs * np.float32(8.0)) * np.float32(0.5)
    sp = _sparse_from_dense_host(h["dims"], m, b, vs, -half)
    obj = np.zeros(1, L.GPU_OBJECT)
    obj["inverse_world"] = np.eye(4, dtype=np.float32).reshape(-1)
    obj["brick_map_dims"] = h["dims"]
    obj["voxel_size"] = vs
    obj["sdf_type"] = L.SDF_TYPE_VOXELIZED
    obj["accumulated_scale"] = (1, 1, 1)
    su = L.scene_uniforms(num_objects=1)
    w, hh = 160, 120
    worst, vox_share = 0.0, 0.0
    for mode in (L.EPS_ABSOLUTE, L.EPS_RELATIVE):
        for eye in ((0.3, 1.0, 1.6), (1.8, 0.6, -0.4), (-1.2, -0.9, 1.1), (0.1, 2.2, 0.2)):
            cu = M.camera_look_at(eye, (0, 0, 0), 60.0, w, hh)
            ref = O.ref_trace(O.RefInputs(b, m, obj, cu, su, L.scene_ext(strategy=L.STRATEGY_BRUTE, hit_eps_mode=mode),
                                          w, hh), prec=64)
            spr = O.sparse_trace(sp, cu, su, L.scene_ext(hit_eps_mode=mode), w, hh, prec=64)
            hit = ref["hit"] == 1
            assert hit.sum() > 2000
            assert np.array_equal(ref["hit"], spr["hit"])
            assert np.array_equal(ref["material_id"][hit], spr["material_id"][hit])
            d = ref["pos"][hit] - np.asarray(eye)
            d /= np.linalg.norm(d, axis=1, keepdims=True)
            cosi = np.abs(np.sum(d * ref["normal"][hit], axis=1))
            eps = 1e-3 * ref["t"][hit] if mode == L.EPS_RELATIVE else np.full(int(hit.sum()), 1e-3)
            along = np.abs(ref["t"][hit] - spr["t"][hit]) * cosi / eps
            worst = max(worst, float(along.max()))
            vox = (ref["voxel_index"][hit] != spr["voxel_index"][hit]) | (ref["brick_slot"][hit] != spr["brick_slot"][hit])
            vox_share = max(vox_share, float(vox.mean()))
    assert worst < 2.0, worst
    assert vox_share < 0.05, vox_share


def test_octant_extents_match_brute_force():
    """orc_sparse_oct against its definition: for a free cell C and octant o,
    E = min over non-free cells M in o's quadrant of max_a |M_a - C_a|, capped
    at 15 (no such cell: 15); E = 0 for non-free cells. Grids with a random
    sprinkle of non-free cells, one cell hugging each face, and the empty grid."""
    rng = np.random.default_rng(4)
    for cells, p_busy in (((9, 7, 5), 0.05), ((20, 3, 18), 0.01), ((4, 4, 4), 0.0), ((17, 17, 17), 0.002)):
        nx, ny, nz = cells
        free = (rng.random(nx * ny * nz) >= p_busy).astype(np.uint8)
        free[0] = 0 if p_busy else free[0]
        got = O.sparse_oct(cells, free)
        idx = np.arange(nx * ny * nz)
        X, Y, Z = idx % nx, (idx // nx) % ny, idx // (nx * ny)
        busy = np.flatnonzero(free == 0)
        for c in rng.choice(idx, size=min(idx.size, 300), replace=False):
            for o in range(8):
                e = (int(got[c]) >> (4 * o)) & 15
                if not free[c]:
                    assert e == 0
                    continue
                s = [(-1 if (o >> a) & 1 else 1) for a in range(3)]
                dx, dy, dz = (X[busy] - X[c]) * s[0], (Y[busy] - Y[c]) * s[1], (Z[busy] - Z[c]) * s[2]
                inq = (dx >= 0) & (dy >= 0) & (dz >= 0)
                want = 15
                if inq.any():
                    want = min(15, int(np.max(np.stack([dx[inq], dy[inq], dz[inq]]), axis=0).min()))
                assert e == want, (cells, c, o, e, want)


def test_brick_extents_are_sound():
    """orc_sparse_bext: in cells that are not entirely free, a brick's octant
    extent E spans a cube of bricks b + [0, E) (in the octant's directions)
    that holds only free bricks (checked by brute force against the brick
    free test); non-free bricks have E = 0; some extents exceed one brick and
    some reach into neighbouring cells."""
    rng = np.random.default_rng(6)
    cells = (3, 2, 3)
    nc = cells[0] * cells[1] * cells[2]
    l2 = np.full(nc, 0xFFFFFFFF, np.uint32)
    occ = [0, 4, 13]
    l1 = np.full(len(occ) * 512, 0xFFFFFFFF, np.uint32)
    nb = 0
    for k, c in enumerate(occ):
        l2[c] = k
        for lb in rng.choice(512, size=6, replace=False):
            l1[k * 512 + lb] = nb
            nb += 1
    sp = O.SparseInputs((0, 0, 0), 0.02, cells, l2, l1, np.zeros(nb * 512, np.uint16), np.zeros(nb * 512, np.uint8))
    NX, NY, NZ = cells[0] * 8, cells[1] * 8, cells[2] * 8
    free = np.zeros((NZ, NY, NX), bool)
    for z in range(NZ):
        for y in range(NY):
            for x in range(NX):
                free[z, y, x] = sp.brick_free(x, y, z)
    ext = sp.bext.reshape(nc, 512)
    checked = big = cross = 0
    for cell in np.flatnonzero(sp.free_l2 == 0):
        cx, cy, cz = cell % cells[0], (cell // cells[0]) % cells[1], cell // (cells[0] * cells[1])
        for lb in range(512):
            b = (cx * 8 + (lb & 7), cy * 8 + ((lb >> 3) & 7), cz * 8 + (lb >> 6))
            for o in range(8):
                e = (int(ext[cell, lb]) >> (4 * o)) & 15
                if not free[b[2], b[1], b[0]]:
                    assert e == 0
                    continue
                assert e >= 1
                lo, hi = [], []
                for a in range(3):
                    if (o >> a) & 1:
                        lo.append(max(b[a] - e + 1, 0)), hi.append(b[a] + 1)
                    else:
                        lo.append(b[a]), hi.append(min(b[a] + e, (NX, NY, NZ)[a]))
                assert free[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]].all(), (cell, lb, o, e)
                checked += 1
                big += e > 1
                cross += any((lo[a] >> 3) != (b[a] >> 3) or ((hi[a] - 1) >> 3) != (b[a] >> 3) for a in range(3))
    assert checked > 10000 and big > 1000 and cross > 100


# ---------------- CSG edits (cpu_apply.rs tests, edit_op.rs tests) ----------------

def _one_brick(fill_d=1.0, fill_mat=0, vs=0.05):
    """A one-cell scene whose brick (0,0,0) is slot 0, filled uniformly (the
    reference tests' setup_pool_with_empty_brick, cpu_apply.rs:303-316)."""
    l2 = np.zeros(1, np.uint32)
    l1 = np.full(512, 0xFFFFFFFF, np.uint32)
    l1[0] = 0
    dist = np.full(512, O.f32_to_f16_bits([fill_d])[0], np.uint16)
    mat = np.full(512, fill_mat, np.uint8)
    return O.SparseInputs((0, 0, 0), vs, (1, 1, 1), l2, l1, dist, mat)


def test_edit_sphere_union_adds_material():
    """cpu_apply.rs:343-363 (sphere_union_adds_material)."""
    sp = _one_brick(1.0, 0)
    op = L.edit_op((0.2, 0.2, 0.2), L.EDIT_SMOOTH_UNION, L.SHAPE_SPHERE, (0.15, 0, 0), 1.0, 0.15 * 0.3,
                   L.FALLOFF_SMOOTH, 5)
    dist, mat, mod = O.sparse_edit(sp, [op])
    assert list(mod) == [0]
    v = 4 + 4 * 8 + 4 * 64
    d = O.f16_bits_to_f32([dist[v]])[0]
    assert d < 1.0
    # closed form at voxel (4,4,4): centre 0.225^3, 0.0433 from the brush centre
    r = np.sqrt(3) * 0.025
    t = r / 0.195
    w = 1 - t * t * (3 - 2 * t)
    eff = (r - 0.15) + (1 - w) * 0.15
    assert abs(d - eff) < 2e-3 and mat[v] == 5
    # voxels outside the falloff radius keep their value and material
    far = 7 + 7 * 8 + 0 * 64
    assert dist[far] == sp.dist[far] or O.f16_bits_to_f32([dist[far]])[0] <= 1.0


def test_edit_sphere_subtract_carves():
    """cpu_apply.rs:366-399 (sphere_subtract_carves)."""
    sp = _one_brick(-0.5, 2)
    op = L.edit_op((0.2, 0.2, 0.2), L.EDIT_SMOOTH_SUBTRACT, L.SHAPE_SPHERE, (0.15, 0, 0), 1.0, 0.05,
                   L.FALLOFF_SMOOTH, 0)
    dist, mat, mod = O.sparse_edit(sp, [op])
    assert len(mod) == 1
    v = 4 + 4 * 8 + 4 * 64
    assert O.f16_bits_to_f32([dist[v]])[0] > -0.5
    assert np.all(mat == 2)


def test_edit_paint_sets_material_near_surface():
    """cpu_apply.rs:402-433 (paint_sets_material_near_surface)."""
    sp = _one_brick(1.0, 0)
    for x in range(8):
        i = x + 4 * 8 + 4 * 64
        sp.dist[i] = O.f32_to_f16_bits([0.01])[0]
        sp.mat[i] = 1
    op = L.edit_op((0.2, 0.2, 0.2), L.EDIT_PAINT, L.SHAPE_SPHERE, (0.5, 0, 0), 1.0, 0.0, L.FALLOFF_SMOOTH, 42)
    dist, mat, mod = O.sparse_edit(sp, [op])
    assert list(mod) == [0]
    assert mat[4 + 4 * 8 + 4 * 64] == 42 & 0x3F
    assert np.array_equal(dist, sp.dist)
    # |d| = 1.0 is not below 2 * 0.5: untouched
    assert mat[0] == 0


def test_edit_outside_the_grid_touches_nothing():
    """cpu_apply.rs:436-441 (no_affected_returns_empty) and edit_op.rs local_aabb."""
    sp = _one_brick(1.0, 0)
    op = L.edit_op((50.0, 50.0, 50.0), L.EDIT_CSG_UNION, L.SHAPE_SPHERE, (0.1, 0, 0), 1.0, 0.0, L.FALLOFF_SMOOTH, 1)
    dist, mat, mod = O.sparse_edit(sp, [op])
    assert len(mod) == 0 and np.array_equal(dist, sp.dist) and np.array_equal(mat, sp.mat)


def test_edit_shape_and_falloff_closed_forms():
    """cpu_apply.rs:446-462: sphere centre -r, box centre -half, smooth falloff 1 at the centre;
    plus the other shapes at points with known distances."""
    lib = O.lib()
    lib.orc_edit_shape.restype = C.c_float
    lib.orc_edit_shape.argtypes = [C.c_uint32, C.c_void_p, C.c_void_p]
    lib.orc_edit_falloff.restype = C.c_float
    lib.orc_edit_falloff.argtypes = [C.c_uint32, C.c_float, C.c_float]

    def shape(s, dims, p):
        d = np.asarray(dims, np.float32)
        q = np.asarray(p, np.float32)
        return lib.orc_edit_shape(s, d.ctypes.data, q.ctypes.data)
    assert abs(shape(L.SHAPE_SPHERE, (1, 0, 0), (0, 0, 0)) + 1) < 1e-6
    assert abs(shape(L.SHAPE_BOX, (0.5, 0.5, 0.5), (0, 0, 0)) + 0.5) < 1e-6
    assert abs(shape(L.SHAPE_BOX, (1, 1, 1), (2, 2, 2)) - np.sqrt(3)) < 1e-6
    assert abs(shape(L.SHAPE_CAPSULE, (0.5, 1, 0), (0, 3, 0)) - 1.5) < 1e-6
    assert abs(shape(L.SHAPE_CYLINDER, (1, 2, 0), (3, 0, 0)) - 2.0) < 1e-6
    assert abs(shape(L.SHAPE_TORUS, (2, 0.5, 0), (2, 0, 0)) + 0.5) < 1e-6
    assert abs(shape(L.SHAPE_PLANE, (0.25, 0, 0), (5, 1, -3)) - 0.75) < 1e-6
    assert abs(lib.orc_edit_falloff(L.FALLOFF_SMOOTH, 0.0, 1.0) - 1.0) < 1e-6
    assert lib.orc_edit_falloff(L.FALLOFF_LINEAR, 1.0, 1.0) == 0.0
    assert abs(lib.orc_edit_falloff(L.FALLOFF_SHARP, 0.5, 1.0) - 0.125) < 1e-7


def test_terrain_height_generator_matches_oracle():
    """The workload generator's vectorised height equals the builder's (oracle) height bit for bit."""
    from rkifield_amd import scenes
    pr, lo, hi = scenes.terrain_scene(n_rocks=0)
    lib = O.lib()
    lib.orc_terrain_height.restype = C.c_float
    lib.orc_terrain_height.argtypes = [C.c_void_p, C.c_float, C.c_float]
    t = np.ascontiguousarray(pr[:1])
    rng = np.random.default_rng(1)
    xs, zs = rng.uniform(-500, 500, 200).astype(np.float32), rng.uniform(-500, 500, 200).astype(np.float32)
    got = scenes.terrain_height(t[0], xs, zs) - np.float32(t[0]["centre"][1])
    want = np.array([lib.orc_terrain_height(t.ctypes.data, float(x), float(z)) for x, z in zip(xs, zs)], np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert np.abs(want).max() < 20 * 1.96875 and np.abs(want).max() > 5


# ---------------- contacts (sdf_collision.rs tests) ----------------

def _ground_plane_scene(vs=0.05, half=(1.6, 1.2, 1.6)):
    """y = 0 ground plane voxelised into a dense grid centred on the origin
    (bricks within 1.8 bricks of the plane, exact per-voxel distances)."""
    bws = vs * 8
    dims = [int(np.ceil(2 * h / bws)) for h in half]
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
                y = origin[1] + by * bws + ((v >> 3) & 7) * vs + vs / 2
                b = np.zeros(512, L.VOXEL_SAMPLE)
                b["word0"] = O.f32_to_f16_bits(y.astype(np.float32)).astype(np.uint32) | (1 << 16)
                bricks.append(b)
    return _sparse_from_dense_host(dims, m, np.concatenate(bricks), np.float32(vs), origin)


def _body(pos, shape, dims, threshold=0.02, rot=(0, 0, 0, 1)):
    b = np.zeros((), L.BODY)
    b["position"], b["shape"], b["dims"], b["threshold"], b["rotation"] = pos, shape, dims, threshold, rot
    return b


def test_contacts_ground_plane_vectors():
    """sdf_collision.rs:484-550: a sphere (r 0.5) above the ground has no
    contacts; at y 0.3 it has contacts with upward normals; at y 0 the deepest
    contact is the bottom pole with penetration ~0.52; contacts come deepest
    first (:646-676)."""
    sp = _ground_plane_scene()
    bodies = np.array([_body((0, 2, 0), L.BODY_SPHERE, (0.5, 0, 0)), _body((0, 0.3, 0), L.BODY_SPHERE, (0.5, 0, 0)),
                       _body((0, 0, 0), L.BODY_SPHERE, (0.5, 0, 0))], L.BODY)
    ct, n = O.sparse_contacts(sp, bodies)
    assert n[0] == 0 and n[1] > 0 and n[2] > 0
    assert np.all(ct[1]["normal"][: n[1], 1] > 0.9)
    assert abs(ct[2]["penetration"][0] - 0.52) < 0.05 and ct[2]["sample"][0] == 3  # bottom pole (0,-r,0)
    pens = ct[2]["penetration"][: n[2]]
    assert np.all(np.diff(pens) <= 0)


def test_contact_sample_points_match_shapes():
    """sdf_collision.rs:578-645: 14 points for a sphere at radius r, 14 for a
    box (8 corners + 6 face centres), 12 for a capsule; with a huge
    threshold every point is a contact whose position is the sample point."""
    sp = _ground_plane_scene()
    big = 1e9
    bodies = np.array([_body((0, 0.2, 0), L.BODY_SPHERE, (0.3, 0, 0), big),
                       _body((0, 0.2, 0), L.BODY_BOX, (0.1, 0.2, 0.3), big),
                       _body((0, 0.2, 0), L.BODY_CAPSULE, (0.25, 0.1, 0), big)], L.BODY)
    ct, n = O.sparse_contacts(sp, bodies)
    assert list(n) == [14, 14, 12]
    rel = [ct[i]["position"][: n[i]] - np.float32([0, 0.2, 0]) for i in range(3)]
    assert np.allclose(np.linalg.norm(rel[0], axis=1), 0.3, atol=1e-6)
    corners = rel[1][np.argsort(ct[1]["sample"][: n[1]])][:8]
    assert np.allclose(np.abs(corners), [0.1, 0.2, 0.3], atol=1e-6)
    assert np.isclose(rel[2][:, 1].max(), 0.35, atol=1e-6) and np.isclose(rel[2][:, 1].min(), -0.35, atol=1e-6)


# ---------------- tile intervals (RKFM_MARCH_TILE_CULL) ----------------

def _camera(eye, target, w, h, jitter=(0.0, 0.0)):
    from rkifield_amd import march as M
    cam = M.camera_look_at(eye, target, 60.0, w, h)
    cam["jitter"][:2] = jitter
    return cam


def _ray_dirs(cam, w, h):
    px, py = np.meshgrid(np.arange(w), np.arange(h))
    ux = (px + 0.5 + cam["jitter"][0]) / w
    uy = (py + 0.5 + cam["jitter"][1]) / h
    nx, ny = ux * 2 - 1, 1 - uy * 2
    f, r, u = (np.asarray(cam[k][:3], np.float64) for k in ("forward", "right", "up"))
    d = f + nx[..., None] * r + ny[..., None] * u
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


def _inside_scene(seed, cells, n_occ, bricks_per_cell, vs):
    """Random occupied cells (boundary cells included), each with a few
    random bricks filled with -vs: a sample any stored brick reaches reads
    below the free value 8 vs."""
    rng = np.random.default_rng(seed)
    nc = cells[0] * cells[1] * cells[2]
    l2 = np.full(nc, L.EMPTY_SLOT, np.uint32)
    occ = rng.choice(nc, size=n_occ, replace=False)
    occ[0] = 0
    occ[1] = nc - 1
    l1 = np.full(n_occ * 512, L.EMPTY_SLOT, np.uint32)
    nb = 0
    for k, c in enumerate(occ):
        l2[c] = k
        lbs = rng.choice(512, size=bricks_per_cell, replace=False)
        lbs[0] = 0 if k % 2 else 511
        for lb in lbs:
            l1[k * 512 + lb] = nb
            nb += 1
    dist = np.full(nb * 512, O.f32_to_f16_bits([-vs])[0], np.uint16)
    return O.SparseInputs((-1.0, -0.5, -2.0), vs, cells, l2, l1, dist, np.zeros(nb * 512, np.uint8))


def test_tile_intervals_are_sound():
    """Outside its tile's [t_lo, t_hi] every sample along every pixel ray
    (step vs/2 up to max_distance) reads at least the free value 8 vs, for
    cameras outside the grid, inside it, inside an occupied cell and with
    sub-pixel jitter; intervals are non-trivial (empty tiles, samples below
    8 vs inside the interval). FP32 and FP64 intervals bracket the same
    samples."""
    vs = 0.05
    cells = (5, 3, 4)
    sp = _inside_scene(3, cells, 9, 6, vs)
    grid = np.asarray(cells, np.float64) * 64 * vs
    origin = np.array([-1.0, -0.5, -2.0])
    occ = np.flatnonzero(sp.l2 != L.EMPTY_SLOT)
    c0 = occ[2]
    inside_occ = origin + (np.array([c0 % cells[0], (c0 // cells[0]) % cells[1], c0 // (cells[0] * cells[1])]) + 0.5) * 64 * vs
    w, h = 56, 40
    su = L.scene_uniforms(num_objects=1, max_distance=40.0)
    views = [((-8.0, 9.0, 14.0), origin + grid / 2, (0.0, 0.0)),
             (tuple(origin + grid * 0.4), tuple(origin + grid), (0.3, -0.4)),
             (tuple(inside_occ), tuple(origin), (-0.5, 0.5)),
             (tuple(origin + grid * np.array([0.5, 1.02, 0.5])), tuple(origin + grid * np.array([0.2, 0.5, 0.9])),
              (0.0, 0.25))]
    total_out = total_in_low = empty = 0
    for eye, tgt, jit in views:
        cam = _camera(eye, tgt, w, h, jit)
        iv = O.sparse_tiles(sp, cam, su, w, h, prec=32)
        iv64 = O.sparse_tiles(sp, cam, su, w, h, prec=64)
        assert np.array_equal(np.isinf(iv[..., 0]), np.isinf(iv64[..., 0]))
        fin = ~np.isinf(iv[..., 0])
        assert np.allclose(iv[fin], iv64[fin], rtol=1e-5, atol=1e-5)
        empty += int((~fin).sum())
        dirs = _ray_dirs(cam, w, h)
        o = np.asarray(cam["position"][:3], np.float64)
        ts = np.arange(0.0, float(su["max_distance"]), vs * 0.5)
        for py in range(0, h, 3):
            for px in range(0, w, 3):
                lo, hi = (float(x) for x in iv[py >> 3, px >> 3])
                for t in ts:
                    d = O.sparse_sample(sp, (o + dirs[py, px] * t).astype(np.float32))
                    if t < lo or t > hi:
                        assert d >= 8 * vs * (1 - 1e-5), (eye, px, py, t, lo, hi, d)
                        total_out += 1
                    elif d < 8 * vs:
                        total_in_low += 1
    assert empty > 0 and total_out > 100000 and total_in_low > 100


def _cross_grid_scene():
    """Copies of test_cross.rkf's 8x4x8 bricks at four places of a 40x16x40
    brick grid (5x2x5 cells): corner and boundary cells, empty space between."""
    data = open(os.path.join(GOLDEN, "test_cross.rkf"), "rb").read()
    h, m, b = O.rkf_load(data)
    cd = [int(x) for x in h["dims"]]
    dims = (40, 16, 40)
    big = np.full(dims[0] * dims[1] * dims[2], L.EMPTY_SLOT, np.uint32)
    bricks = []
    for k, (ox, oy, oz) in enumerate(((0, 0, 0), (16, 8, 16), (32, 4, 8), (8, 12, 32))):
        for bz in range(cd[2]):
            for by in range(cd[1]):
                for bx in range(cd[0]):
                    e = m[bx + by * cd[0] + bz * cd[0] * cd[1]]
                    if e == L.EMPTY_SLOT:
                        continue
                    big[(ox + bx) + (oy + by) * dims[0] + (oz + bz) * dims[0] * dims[1]] = len(bricks)
                    bricks.append(b[e * 512:(e + 1) * 512])
    vs = np.float32(h["voxel_size"])
    return _sparse_from_dense_host(dims, big, np.concatenate(bricks), vs, (-1.0, -0.5, -1.5)), float(vs)


def test_tile_culled_march_agrees_with_unculled():
    """The culled march (FP32 twin and FP64 oracle) finds the same surfaces
    as the plain march: same hit mask, depths within the hit band (the
    relative hit threshold), same materials, fewer loop iterations."""
    sp, vs = _cross_grid_scene()
    w, h = 64, 48
    su = L.scene_uniforms(num_objects=1, max_distance=30.0)
    for eye, tgt in (((3.5, 2.2, 5.5), (1.5, 0.8, 1.5)), ((0.3, 0.2, 1.0), (-1.0, 0.0, -1.0))):
        cam = _camera(eye, tgt, w, h)
        plain = L.scene_ext(hit_eps_mode=L.EPS_RELATIVE)
        cull = L.scene_ext(hit_eps_mode=L.EPS_RELATIVE, march_flags=L.MARCH_TILE_CULL)
        for prec in (32, 64):
            a = O.sparse_trace(sp, cam, su, plain, w, h, prec=prec)
            b = O.sparse_trace(sp, cam, su, cull, w, h, prec=prec)
            assert a["hit"].sum() > 200
            assert np.array_equal(a["hit"], b["hit"])
            hit = a["hit"] == 1
            # a march stops anywhere its sample drops below the hit threshold
            # (relative: 1e-3 t), so two step sequences agree to that band
            rel = np.abs(a["t"][hit] - b["t"][hit]) / a["t"][hit]
            assert rel.max() <= float(su["hit_threshold"]) * 1.01 and np.median(rel) <= 1e-6
            assert np.array_equal(a["material_id"], b["material_id"])
            assert b["steps"].sum() < a["steps"].sum()

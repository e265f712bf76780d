"""GPU parity of the sparse two-level scene path against the oracle.

* GPU voxeliser: every stored voxel distance and every allocation decision
  equals the brute-force oracle field (all primitives, all groups), bit for bit.
* Free-space masks: equal to the oracle's dense masks.
* March (plain, tuned and split variants), shadow rays and point queries: bit-exact
  vs the FP32 twin; depth/normal within tolerance of the FP64 oracle.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle_py as O
from rkifield_amd import layouts as L
from rkifield_amd import march as M
from rkifield_amd import scenes

pytestmark = pytest.mark.gpu

DEPTH_REL_TOL = 1e-4
NORMAL_RAD_TOL = 1e-3


def _say(msg):
    print(msg, flush=True)


def _mix64(z):
    z = (z + np.uint64(0x9E3779B97F4A7C15))
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def _small_prims(seed=7, n=48, half=3.0):
    p = scenes.primitive_field(seed=seed, n=n, half_cube=half, size_range=(0.3, 1.5), group=4)
    return p


def _oracle_inputs(sc):
    i = sc.info
    l2, l1, dist, mat = sc.download()
    return O.SparseInputs(i["origin"], float(i["voxel_size"]), i["cells"], l2, l1, dist, mat), (l2, l1, dist, mat)


@pytest.fixture(scope="module")
def small_scene(ctx):
    prims = _small_prims()
    lo, hi = scenes.prim_bounds(prims, 0.5)
    _say("case: sparse build, 48 prims, vs 0.05")
    with np.errstate(over="ignore"):
        sc = M.SparseScene.build(ctx, prims, 0.05, lo, hi, smin_k=0.1, mat_seed=scenes.material_seed(7))
    yield prims, sc, lo
    sc.destroy()


def test_voxeliser_matches_brute_force_field(small_scene):
    prims, sc, lo = small_scene
    sp, (l2, l1, dist, mat) = _oracle_inputs(sc)
    i = sc.info
    assert i["n_bricks"] > 100
    vs = np.float32(i["voxel_size"])
    bws = vs * np.float32(8)
    cells = [int(c) for c in i["cells"]]
    lib = O.lib()
    lib.orc_scene_field.restype = C.c_float
    lib.orc_scene_field.argtypes = [C.c_void_p, C.c_uint32, C.c_float, C.c_void_p]
    pr = np.ascontiguousarray(prims)
    field = lambda p: lib.orc_scene_field(pr.ctypes.data, pr.size, np.float32(0.1), np.asarray(p, np.float32).ctypes.data)
    origin = np.asarray(i["origin"], np.float32)
    band = bws * np.float32(1.8)
    nb = [c * 8 for c in cells]
    rng = np.random.default_rng(0)
    # allocation decisions: every brick of every occupied cell plus random bricks
    checked = 0
    for gbz in range(nb[2]):
        for gby in range(nb[1]):
            for gbx in range(nb[0]):
                cell = (gbx >> 3) + (gby >> 3) * cells[0] + (gbz >> 3) * cells[0] * cells[1]
                blk = l2[cell]
                slot = L.EMPTY_SLOT if blk == L.EMPTY_SLOT else l1[blk * 512 + (gbx & 7) + (gby & 7) * 8 + (gbz & 7) * 64]
                if blk == L.EMPTY_SLOT and rng.random() > 0.02:
                    continue
                hb = bws * np.float32(0.5)
                c = [(origin[0] + np.float32(gbx) * bws) + hb, (origin[1] + np.float32(gby) * bws) + hb,
                     (origin[2] + np.float32(gbz) * bws) + hb]
                need = abs(field(c)) < band
                assert need == (slot != L.EMPTY_SLOT), (gbx, gby, gbz)
                checked += 1
                if slot != L.EMPTY_SLOT and rng.random() < 0.05:
                    half = vs * np.float32(0.5)
                    for v in rng.integers(0, 512, 16):
                        vx, vy, vz = v & 7, (v >> 3) & 7, v >> 6
                        p = [(origin[0] + np.float32(gbx) * bws) + (np.float32(vx) * vs + half),
                             (origin[1] + np.float32(gby) * bws) + (np.float32(vy) * vs + half),
                             (origin[2] + np.float32(gbz) * bws) + (np.float32(vz) * vs + half)]
                        want = O.f32_to_f16_bits([field(p)])[0]
                        assert dist[slot * 512 + v] == want
                        key = np.uint64((gbx * 8 + vx) | ((gby * 8 + vy) << 21) | ((gbz * 8 + vz) << 42))
                        seed = np.uint64(scenes.material_seed(7))
                        with np.errstate(over="ignore"):
                            m = int(_mix64(seed ^ _mix64(key)) & np.uint64(15))
                        assert mat[slot * 512 + v] == m
    assert checked > 500


def test_free_masks_match_oracle(small_scene):
    prims, sc, lo = small_scene
    sp, _ = _oracle_inputs(sc)
    fidx, bits, oct_, bext = sc.download_free()
    assert np.array_equal(oct_, sp.oct)
    assert (oct_ != 0).any() and (oct_ == 0).any()
    # brick extents of every row cell, all 8 octant planes
    ref = sp.bext.reshape(-1, 512)
    rows = np.flatnonzero(fidx != 0xFFFFFFFF)
    planes = bext.reshape(-1, 8, 256)
    for cell in rows:
        pl = planes[fidx[cell]]
        got = np.zeros(512, np.uint32)
        for o in range(8):
            nib = np.empty(512, np.uint32)
            nib[0::2] = pl[o] & 15
            nib[1::2] = pl[o] >> 4
            got |= nib << np.uint32(4 * o)
        assert np.array_equal(got, ref[cell]), cell
    assert (ref[rows] != 0).any()
    cells = [int(c) for c in sc.info["cells"]]
    nb = [c * 8 for c in cells]
    for cz in range(cells[2]):
        for cy in range(cells[1]):
            for cx in range(cells[0]):
                cell = cx + cy * cells[0] + cz * cells[0] * cells[1]
                assert (fidx[cell] == 0xFFFFFFFF) == bool(sp.free_l2[cell])
                if fidx[cell] == 0xFFFFFFFF:
                    continue
                row = bits[fidx[cell] * 8: fidx[cell] * 8 + 8]
                for lb in range(512):
                    bx, by, bz = cx * 8 + (lb & 7), cy * 8 + ((lb >> 3) & 7), cz * 8 + (lb >> 6)
                    notfree = (int(row[lb >> 6]) >> (lb & 63)) & 1
                    assert notfree == (0 if sp.brick_free(bx, by, bz) else 1)


def test_tile_intervals_match_oracle(ctx, small_scene):
    """rkfm_sparse_tile_intervals equals the oracle's FP32 intervals bit for
    bit: camera outside the grid, inside it, inside an occupied cell, with
    jitter, odd image sizes."""
    prims, sc, lo = small_scene
    sp, _ = _oracle_inputs(sc)
    l2 = sp.l2
    cells = [int(c) for c in sc.info["cells"]]
    c0 = int(np.flatnonzero(l2 != L.EMPTY_SLOT)[len(np.flatnonzero(l2 != L.EMPTY_SLOT)) // 2])
    cw = float(sc.info["voxel_size"]) * 64
    org = np.asarray(sc.info["origin"], np.float64)
    in_occ = org + (np.array([c0 % cells[0], (c0 // cells[0]) % cells[1], c0 // (cells[0] * cells[1])]) + 0.5) * cw
    su = L.scene_uniforms(num_objects=1, max_distance=50.0)
    views = [((6.0, 3.0, 7.0), (0, 0, 0), 96, 64, (0.0, 0.0)), ((0.5, 0.2, -0.3), (2.0, 1.0, 3.0), 101, 37, (0.3, -0.2)),
             (tuple(in_occ), (0.0, 0.0, 0.0), 64, 64, (-0.5, 0.5)),
             ((20.0, 10.0, 24.0), (40.0, 20.0, 48.0), 80, 48, (0.1, 0.1))]
    nonempty = empty = 0
    for eye, tgt, w, h, jit in views:
        cam = M.camera_look_at(eye, tgt, 60.0, w, h)
        cam["jitter"][:2] = jit
        _say(f"case: tile intervals {w}x{h} eye {eye}")
        got = sc.tile_intervals(cam, su, w, h)
        want = O.sparse_tiles(sp, cam, su, w, h, prec=32)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
        fin = np.isfinite(got[..., 0])
        nonempty += int(fin.sum())
        empty += int((~fin).sum())
    assert nonempty > 50 and empty > 10


@pytest.mark.parametrize("variant,cull", [(0, 0), (1, 0), (2, 0), (0, 1), (1, 1), (2, 1)])
def test_sparse_march_parity(ctx, small_scene, variant, cull):
    prims, sc, lo = small_scene
    sp, _ = _oracle_inputs(sc)
    w, h = 96, 64
    cam = M.camera_look_at((6.0, 3.0, 7.0), (0, 0, 0), 60.0, w, h)
    su = L.scene_uniforms(num_objects=1)
    ext = L.scene_ext(hit_eps_mode=L.EPS_RELATIVE, shadow_enable=1,
                      light_dir=(np.array([0.5, 1.0, 0.3]) / np.linalg.norm([0.5, 1.0, 0.3])),
                      march_flags=L.MARCH_TILE_CULL if cull else 0)
    _say(f"case: sparse march variant {variant} cull {cull} {w}x{h}")
    gb = M.GBuffer(ctx, w, h, debug=True, shadow=True)
    sc.march(cam, su, ext, gb, variant=variant)
    gpu = gb.read()
    res32 = O.sparse_trace(sp, cam, su, ext, w, h, prec=32)
    assert res32["hit"].sum() > 300
    pos, nrm, mat, mot = O.pack_gbuffer(res32)
    hit = res32["hit"] == 1
    assert np.array_equal(gpu["position"][:, 3] != L.MAX_FLOAT, hit)
    assert np.array_equal(gpu["material"], mat)
    for f in ("brick_slot", "voxel_index", "cell_index", "steps", "evals", "shadow_steps", "shadow_evals"):
        assert np.array_equal(gpu["debug"][f], res32[f]), f
    assert np.array_equal(gpu["position"].view(np.uint32), pos.view(np.uint32))
    assert np.array_equal(gpu["normal"], nrm)
    assert np.array_equal(gpu["motion"].view(np.uint32), mot.view(np.uint32))
    assert np.array_equal(gpu["shadow"].view(np.uint32), res32["shadow"].astype(np.float32).view(np.uint32))
    res64 = O.sparse_trace(sp, cam, su, ext, w, h, prec=64)
    same = (res64["hit"] == res32["hit"]) & (res64["steps"] == res32["steps"]) & (res64["path"] == res32["path"])
    both = same & hit
    rel = np.abs(gpu["position"][both, 3] - res64["t"][both]) / res64["t"][both]
    n = gpu["debug"]["normal"][both].astype(np.float64)
    ang = np.arccos(np.clip(np.sum(n * res64["normal"][both], axis=1), -1, 1))
    _say(f"sparse: same-path {same.mean():.4f} max rel depth {rel.max():.3e} max normal angle {ang.max():.3e}")
    assert same.mean() > 0.99
    assert rel.max() <= DEPTH_REL_TOL
    assert ang.max() <= NORMAL_RAD_TOL


@pytest.mark.parametrize("cull", [0, 1])
def test_sparse_march_ragged_view(ctx, small_scene, cull):
    """A view whose edge tiles are partly outside it: those wavefronts hand
    their last rays to 8-lane groups with lanes that hold no pixel. Every
    field bit-exact against the FP32 twin."""
    prims, sc, lo = small_scene
    sp, _ = _oracle_inputs(sc)
    w, h = 61, 45
    cam = M.camera_look_at((5.5, 2.5, -6.5), (0, 0, 0), 60.0, w, h)
    su = L.scene_uniforms(num_objects=1)
    ext = L.scene_ext(hit_eps_mode=L.EPS_RELATIVE, shadow_enable=1,
                      light_dir=(np.array([0.5, 1.0, 0.3]) / np.linalg.norm([0.5, 1.0, 0.3])),
                      march_flags=L.MARCH_TILE_CULL if cull else 0)
    _say(f"case: sparse march ragged {w}x{h} cull {cull}")
    gb = M.GBuffer(ctx, w, h, debug=True, shadow=True)
    sc.march(cam, su, ext, gb)
    gpu = gb.read()
    res32 = O.sparse_trace(sp, cam, su, ext, w, h, prec=32)
    assert res32["hit"].sum() > 200
    edge = (np.arange(w * h) % w >= 56) | (np.arange(w * h) // w >= 40)
    assert res32["hit"][edge].sum() > 10
    pos, nrm, mat, mot = O.pack_gbuffer(res32)
    for f in ("brick_slot", "voxel_index", "cell_index", "steps", "evals", "shadow_steps", "shadow_evals"):
        assert np.array_equal(gpu["debug"][f], res32[f]), f
    assert np.array_equal(gpu["position"].view(np.uint32), pos.view(np.uint32))
    assert np.array_equal(gpu["normal"], nrm)
    assert np.array_equal(gpu["material"], mat)
    assert np.array_equal(gpu["motion"].view(np.uint32), mot.view(np.uint32))
    assert np.array_equal(gpu["shadow"].view(np.uint32), res32["shadow"].astype(np.float32).view(np.uint32))


def test_tile_order_matches_oracle(ctx, small_scene):
    """rkfm_sparse_tile_order (the culled march's launch order): per-tile
    cell counts equal the oracle's FP32 binning, the order is a permutation
    of the tiles by descending min(count, 255); frames above 65,536 tiles are
    refused."""
    prims, sc, lo = small_scene
    sp, _ = _oracle_inputs(sc)
    su = L.scene_uniforms(num_objects=1)
    for eye, w, h in (((6.0, 3.0, 7.0), 96, 64), ((5.5, 2.5, -6.5), 61, 45)):
        _say(f"case: tile order {w}x{h}")
        cam = M.camera_look_at(eye, (0, 0, 0), 60.0, w, h)
        counts, order = sc.tile_order(cam, su, w, h)
        assert np.array_equal(counts, O.sparse_tile_counts(sp, cam, su, w, h))
        assert len(np.unique(counts)) > 3
        assert np.array_equal(np.sort(order), np.arange(counts.size))
        key = np.minimum(counts.ravel()[order], 255).astype(np.int64)
        assert np.all(np.diff(key) <= 0)
    cam = M.camera_look_at((6.0, 3.0, 7.0), (0, 0, 0), 60.0, 4096, 4096)
    with pytest.raises(M.RkfMarchError):
        sc.tile_order(cam, su, 4096, 4096)


def test_point_queries_parity(ctx, small_scene):
    prims, sc, lo = small_scene
    sp, _ = _oracle_inputs(sc)
    rng = np.random.default_rng(3)
    n = 5000
    hi = lo + np.asarray(sc.info["cells"], np.float32) * np.float32(sc.info["voxel_size"] * 64)
    pts = rng.uniform(lo - 1, hi + 1, (n, 3)).astype(np.float32)
    _say("case: point queries 5000")
    d_in = ctx.to_device(pts)
    d_out = ctx.array((n,), L.POINT_RESULT)
    sc.query(d_in, d_out, n)
    got = d_out.download()
    want = O.sparse_query(sp, pts, prec=32)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


def test_from_dense_matches_reference_layout(ctx):
    """test_cross.rkf streamed into a resident sparse scene at the origin,
    by the oracle's reader + from_dense and by the product's .rkf loader."""
    import os
    data = open(os.path.join(os.path.dirname(__file__), "golden", "test_cross.rkf"), "rb").read()
    h, m, b = O.rkf_load(data)
    vs = np.float32(h["voxel_size"])
    half = h["dims"].astype(np.float32) * vs * np.float32(8) * np.float32(0.5)
    _say("case: from_dense test_cross")
    sc = M.SparseScene.from_dense(ctx, h["dims"], m, b, float(vs), -half)
    sp, _ = _oracle_inputs(sc)
    w, hh = 64, 48
    cam = M.camera_look_at((0.3, 1.0, 1.6), (0, 0, 0), 60.0, w, hh)
    su = L.scene_uniforms(num_objects=1)
    ext = L.scene_ext()
    gb = M.GBuffer(ctx, w, hh, debug=True)
    sc.march(cam, su, ext, gb, variant=1)
    gpu = gb.read()
    res = O.sparse_trace(sp, cam, su, ext, w, hh, prec=32)
    pos, nrm, mat, mot = O.pack_gbuffer(res)
    assert res["hit"].sum() > 200
    assert np.array_equal(gpu["material"], mat)
    assert np.array_equal(gpu["position"].view(np.uint32), pos.view(np.uint32))
    assert np.array_equal(gpu["debug"]["brick_slot"], res["brick_slot"])
    # the same asset through the product's .rkf loader (centred grid by
    # default): identical index, pool and march output
    _say("case: from_rkf test_cross")
    sr = M.SparseScene.from_rkf(ctx, data)
    assert np.array_equal(sr.info["origin"], sc.info["origin"])
    for x, y in zip(sr.download(), sc.download()):
        assert np.array_equal(x, y)
    gb2 = M.GBuffer(ctx, w, hh, debug=True)
    sr.march(cam, su, ext, gb2, variant=0)
    got = gb2.read()
    for k in ("position", "normal", "material", "motion", "debug"):
        assert np.array_equal(got[k].view(np.uint8), gpu[k].view(np.uint8)), k
    sr.destroy()
    sc.destroy()


def test_sparse_errors(ctx):
    prims = _small_prims(n=4)
    lo, hi = scenes.prim_bounds(prims, 0.5)
    with pytest.raises(M.RkfMarchError) as e:
        M.SparseScene.build(ctx, prims, 0.05, lo, hi, max_bricks=1)
    assert e.value.code == M.ERR["OOM"]
    with pytest.raises(M.RkfMarchError):
        M.SparseScene.build(ctx, prims, -1.0, lo, hi)
    # brick grids are limited to 2^18 bricks per axis (packed per-ray cache keys)
    with pytest.raises(M.RkfMarchError) as e:
        M.SparseScene.from_dense(ctx, (8 * 32768 + 8, 1, 1), np.full(8 * 32768 + 8, 0xFFFFFFFF, np.uint32),
                                 np.zeros(512, L.VOXEL_SAMPLE), 0.05, (0.0, 0.0, 0.0))
    assert e.value.code == M.ERR["INVALID_ARG"]


def test_sphere_casts_match_oracle(ctx, small_scene):
    """rkfm_sparse_sphere_cast: every result field bit-exact against the FP32
    twin; radii from below a voxel to above the 8 vs field cap, casts from
    outside and inside the grid, hits and misses."""
    prims, sc, lo = small_scene
    sp, _ = _oracle_inputs(sc)
    rng = np.random.default_rng(8)
    n = 6000
    casts = np.zeros(n, L.SPHERE_CAST)
    casts["origin"] = rng.uniform(-5.0, 5.0, (n, 3))
    d = rng.normal(size=(n, 3))
    casts["dir"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    casts["radius"] = rng.uniform(0.01, 0.8, n)
    casts["max_t"] = rng.uniform(1.0, 12.0, n)
    _say(f"case: sphere casts {n}")
    cd, rd = ctx.to_device(casts), ctx.array((n,), L.CAST_RESULT)
    sc.sphere_cast(cd, rd, n)
    got = rd.download()
    want = O.sparse_cast(sp, casts)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))
    assert 500 < got["hit"].sum() < n - 500
    assert (got["steps"] > 5).sum() > 500


def _check_view(gpu, res32):
    pos, nrm, mat, mot = O.pack_gbuffer(res32)
    for f in ("brick_slot", "voxel_index", "cell_index", "steps", "evals", "shadow_steps", "shadow_evals"):
        assert np.array_equal(gpu["debug"][f], res32[f]), f
    assert np.array_equal(gpu["position"].view(np.uint32), pos.view(np.uint32))
    assert np.array_equal(gpu["normal"], nrm)
    assert np.array_equal(gpu["material"], mat)
    assert np.array_equal(gpu["motion"].view(np.uint32), mot.view(np.uint32))
    assert np.array_equal(gpu["shadow"].view(np.uint32), res32["shadow"].astype(np.float32).view(np.uint32))


def test_culled_views_back_to_back(ctx, small_scene):
    """Culled views marched one after another bin into the tile buffers the
    previous plain march put back to the cleared state, with no clearing
    launch: repeated views, a smaller view after a larger one, a larger one
    after a smaller one, and the tile-order and tile-interval entry points
    (which leave those buffers in use) in between -- every view bit-exact
    against the FP32 twin, and the tile counts right after a march equal to
    the oracle's."""
    prims, sc, lo = small_scene
    sp, _ = _oracle_inputs(sc)
    su = L.scene_uniforms(num_objects=1)
    ld = np.array([0.5, 1.0, 0.3]) / np.linalg.norm([0.5, 1.0, 0.3])
    ext = L.scene_ext(hit_eps_mode=L.EPS_RELATIVE, shadow_enable=1, light_dir=ld, march_flags=L.MARCH_TILE_CULL)
    a, b, c = ((6.0, 3.0, 7.0), 96, 64), ((5.5, 2.5, -6.5), 61, 45), ((-6.0, 4.0, 5.0), 40, 24)
    steps = [(a, None), (a, None), (b, None), (b, "order"), (b, None), (c, "intervals"), (c, None), (c, None), (a, None)]
    for k, ((eye, w, h), between) in enumerate(steps):
        cam = M.camera_look_at(eye, (0, 0, 0), 60.0, w, h)
        if between == "order":
            counts, _ = sc.tile_order(cam, su, w, h)  # right after a march of this view
            assert np.array_equal(counts, O.sparse_tile_counts(sp, cam, su, w, h))
        elif between == "intervals":
            sc.tile_intervals(cam, su, w, h)
        _say(f"case: back-to-back culled view {k}: {w}x{h} ({between or 'march only'} before)")
        gb = M.GBuffer(ctx, w, h, debug=True, shadow=True)
        sc.march(cam, su, ext, gb)
        res32 = O.sparse_trace(sp, cam, su, ext, w, h, prec=32)
        assert res32["hit"].sum() > 50
        _check_view(gb.read(), res32)

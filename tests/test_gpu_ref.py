"""GPU parity of the drop-in kernel (reference layouts) against the oracle.

Bit-exact: hit mask, packed material (material | secondary | blend | object id),
brick slot, voxel index, object index, step and evaluation counts, and (same
operation order) every float of the G-buffer vs the FP32 twin.
Tolerances vs the FP64 oracle (rays whose discrete path is the same):
depth <= 1e-4 relative, normal <= 1e-3 rad (f32 debug normal).
"""
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


def _upload(ctx, sc, strategy=L.STRATEGY_AUTO, tiles=True, coarse=True):
    gs = M.GpuScene(ctx)
    gs.upload_brick_pool(sc.pool)
    gs.upload_brick_maps(sc.maps)
    gs.upload_objects(sc.objects)
    gs.upload_bvh(M.build_bvh(sc.objects))
    gs.update_camera(sc.cam)
    ext = sc.ext.copy()
    ext["strategy"] = strategy
    gs.update_scene(sc.su, ext)
    tc = M.TileObjectCullPass(ctx, sc.width, sc.height)
    cf = M.CoarseField.from_scene_aabbs(ctx, sc.aabbs())
    if tiles:
        tc.dispatch()
    else:
        tc.clear()
    if coarse:
        cf.populate_and_upload()
    else:
        cf.clear()
    return gs, tc, cf, ext


def _compare(gpu, res, width, height):
    pos, nrm, mat, mot = O.pack_gbuffer(res)
    dbg = gpu["debug"]
    hit = res["hit"] == 1
    ghit = gpu["position"][:, 3] != L.MAX_FLOAT
    assert np.array_equal(ghit, hit), f"hit mask differs on {np.sum(ghit != hit)} rays"
    assert np.array_equal(gpu["material"], mat)
    assert np.array_equal(dbg["brick_slot"], res["brick_slot"])
    assert np.array_equal(dbg["voxel_index"], res["voxel_index"])
    assert np.array_equal(dbg["cell_index"], res["cell_index"])
    assert np.array_equal(dbg["steps"], res["steps"])
    assert np.array_equal(dbg["evals"], res["evals"])
    assert np.array_equal(gpu["position"].view(np.uint32), pos.view(np.uint32))
    assert np.array_equal(gpu["normal"], nrm)
    assert np.array_equal(gpu["motion"].view(np.uint32), mot.view(np.uint32))
    dn = np.asarray(res["normal"], np.float32)
    assert np.array_equal(dbg["normal"][hit].view(np.uint32), dn[hit].view(np.uint32))


def _tolerance(gpu, res64, res32):
    same = (res64["hit"] == res32["hit"]) & (res64["steps"] == res32["steps"]) & (res64["path"] == res32["path"])
    hit = same & (res32["hit"] == 1)
    t = gpu["position"][:, 3].astype(np.float64)
    rel = np.abs(t[hit] - res64["t"][hit]) / res64["t"][hit]
    n = gpu["debug"]["normal"].astype(np.float64)
    cosang = np.clip(np.sum(n[hit] * res64["normal"][hit], axis=1), -1.0, 1.0)
    ang = np.arccos(cosang)
    return same.mean(), rel.max(initial=0.0), ang.max(initial=0.0)


@pytest.fixture(scope="module")
def spheres():
    return scenes.spheres_scene(seed=0)


def test_producers_match_oracle(ctx, spheres):
    sc = spheres
    _say("case: producers, spheres seed 0, 640x360")
    gs, tc, cf, ext = _upload(ctx, sc)
    idx, counts = tc.read()
    oidx, ocounts = O.tile_cull(sc.cam, sc.su, sc.objects, sc.width, sc.height)
    assert np.array_equal(counts, ocounts)
    for t in range(counts.size):
        c = counts[t]
        assert np.array_equal(idx[t * 32:t * 32 + c], oidx[t * 32:t * 32 + c])
    cells, u = cf.read()
    ocells = O.coarse_populate(cf.origin, cf.dims, cf.voxel_size, sc.objects)
    assert np.array_equal(cells, ocells)
    assert np.allclose(u["origin_cam_rel"][:3], cf.origin - sc.cam["position"][:3])


@pytest.mark.parametrize("strategy", [L.STRATEGY_TILED])
def test_config0_full_parity(ctx, spheres, strategy):
    sc = spheres
    _say(f"case: config0 640x360 strategy {strategy}")
    gs, tc, cf, ext = _upload(ctx, sc, strategy)
    gb = M.GBuffer(ctx, sc.width, sc.height, debug=True)
    M.RayMarchPass(ctx).dispatch(gs, gb, tc, cf)
    gpu = gb.read()
    idx, counts = tc.read()
    cells, u = cf.read()
    inp = O.RefInputs(sc.pool, sc.maps, sc.objects, sc.cam, sc.su, ext, sc.width, sc.height,
                      tile_idx=idx, tile_counts=counts, coarse=cells, coarse_u=u)
    res32, _ = O.ref_trace_parallel(inp)
    assert res32["hit"].sum() > 1000
    _compare(gpu, res32, sc.width, sc.height)
    res64 = O.ref_trace(inp, prec=64)
    same, rel, ang = _tolerance(gpu, res64, res32)
    _say(f"config0: same-path {same:.5f} max rel depth {rel:.3e} max normal angle {ang:.3e} rad")
    assert same > 0.995
    assert rel <= DEPTH_REL_TOL
    assert ang <= NORMAL_RAD_TOL


@pytest.mark.parametrize("strategy", [L.STRATEGY_BVH, L.STRATEGY_BRUTE])
def test_fallback_strategies_parity(ctx, strategy):
    sc = scenes.spheres_scene(seed=3, n=12, width=96, height=64, half_cube=3.0, cam_dist=9.0)
    _say(f"case: fallback strategy {strategy} 96x64")
    gs, tc, cf, ext = _upload(ctx, sc, strategy, tiles=False, coarse=False)
    gb = M.GBuffer(ctx, sc.width, sc.height, debug=True)
    M.RayMarchPass(ctx).dispatch(gs, gb)
    gpu = gb.read()
    nodes = M.build_bvh(sc.objects)
    inp = O.RefInputs(sc.pool, sc.maps, sc.objects, sc.cam, sc.su, ext, sc.width, sc.height,
                      bvh=nodes)
    res32 = O.ref_trace(inp, prec=32)
    assert res32["hit"].sum() > 50
    _compare(gpu, res32, sc.width, sc.height)


def test_ragged_sizes_and_absolute_eps(ctx):
    """Image sizes that are not multiples of 8 or 16; absolute hit threshold."""
    sc = scenes.spheres_scene(seed=5, n=10, width=37, height=23, half_cube=2.0, cam_dist=7.0, rel_eps=False)
    _say("case: ragged 37x23 abs eps")
    gs, tc, cf, ext = _upload(ctx, sc)
    gb = M.GBuffer(ctx, sc.width, sc.height, debug=True)
    M.RayMarchPass(ctx).dispatch(gs, gb, tc, cf)
    gpu = gb.read()
    idx, counts = tc.read()
    cells, u = cf.read()
    inp = O.RefInputs(sc.pool, sc.maps, sc.objects, sc.cam, sc.su, ext, sc.width, sc.height,
                      tile_idx=idx, tile_counts=counts, coarse=cells, coarse_u=u)
    _compare(gpu, O.ref_trace(inp, prec=32), sc.width, sc.height)


def test_empty_scene_and_errors(ctx):
    sc = scenes.spheres_scene(seed=1, n=4, width=16, height=16, half_cube=2.0, cam_dist=6.0)
    _say("case: errors")
    gs, tc, cf, ext = _upload(ctx, sc)
    su = sc.su.copy()
    su["num_objects"] = 99
    gs.update_scene(su, ext)
    gb = M.GBuffer(ctx, 16, 16)
    with pytest.raises(M.RkfMarchError) as e:
        M.RayMarchPass(ctx).dispatch(gs, gb, tc, cf)
    assert e.value.code == M.ERR["OUT_OF_RANGE"]
    su["num_objects"] = 0
    gs.update_scene(su, ext)
    tc.dispatch()
    M.RayMarchPass(ctx).dispatch(gs, gb, tc, cf)
    out = gb.read()
    assert np.all(out["position"][:, 3] == L.MAX_FLOAT) and np.all(out["material"] == 0)
    gb2 = M.GBuffer(ctx, 32, 16)
    su["num_objects"] = 4
    gs.update_scene(su, ext)
    with pytest.raises(M.RkfMarchError) as e:
        M.RayMarchPass(ctx).dispatch(gs, gb2, tc, cf)  # tile grid built for 16x16
    assert e.value.code == M.ERR["LAYOUT"]


def test_upload_index_limits(ctx):
    """Voxels and map entries are indexed in 32 bits (the reference's u32
    indices): uploads past slot 2^23 - 2 or map entry 2^32 - 2 are refused
    before any device memory is touched; the last allowed ones are taken."""
    _say("case: upload limits")
    gs = M.GpuScene(ctx)
    brick = np.zeros(512, L.VOXEL_SAMPLE)
    for first in ((1 << 23) - 1, 1 << 23, 1 << 40):
        with pytest.raises(M.RkfMarchError) as e:
            gs.upload_brick_pool(brick, first_slot=first)
        assert e.value.code == M.ERR["OUT_OF_RANGE"]
    for off in ((1 << 32) - 1, 1 << 32):
        with pytest.raises(M.RkfMarchError) as e:
            gs.upload_brick_maps(np.zeros(1, np.uint32), offset=off)
        assert e.value.code == M.ERR["OUT_OF_RANGE"]
    gs.upload_brick_pool(brick, first_slot=0)
    gs.upload_brick_maps(np.zeros(4, np.uint32), offset=0)


def test_host_tile_lists(ctx):
    """Tile lists made on the host and handed over with rkfm_set_tiles (which
    also makes the launch order from them): the oracle's cull, then lists with
    each tile's objects reversed, about a third of the tiles emptied and one
    tile naming 40 objects (past the 32-entry capacity: the first 32 are
    used) -- each bit-exact against the FP32 twin given the same lists."""
    sc = scenes.spheres_scene(seed=2, n=40, width=96, height=64, half_cube=4.0, cam_dist=12.0)
    _say("case: host tile lists 96x64, 40 spheres")
    gs, tc, cf, ext = _upload(ctx, sc)
    idx, counts = O.tile_cull(sc.cam, sc.su, sc.objects, sc.width, sc.height)
    assert counts.max() > 8 and (counts == 1).any() and (counts == 0).any()
    lists = idx.reshape(-1, 32).copy()
    c2 = counts.copy()
    for t in range(c2.size):
        lists[t, :c2[t]] = lists[t, :c2[t]][::-1].copy()
    rng = np.random.default_rng(7)
    c2[rng.random(c2.size) < 0.3] = 0
    full = int(np.argmax(counts))
    lists[full] = np.arange(32, dtype=np.uint32)
    c2[full] = 40
    cells, u = cf.read()
    for name, ti, tcn in (("oracle cull", idx, counts), ("reversed, emptied, over capacity", lists.reshape(-1), c2)):
        _say(f"case: host tile lists: {name}")
        tc.set_lists(ti, tcn)
        gb = M.GBuffer(ctx, sc.width, sc.height, debug=True)
        M.RayMarchPass(ctx).dispatch(gs, gb, tc, cf)
        inp = O.RefInputs(sc.pool, sc.maps, sc.objects, sc.cam, sc.su, ext, sc.width, sc.height,
                          tile_idx=np.ascontiguousarray(ti, np.uint32), tile_counts=np.ascontiguousarray(tcn, np.uint32),
                          coarse=cells, coarse_u=u)
        res = O.ref_trace(inp, prec=32)
        assert res["hit"].sum() > 100
        _compare(gb.read(), res, sc.width, sc.height)


def test_host_tile_lists_every_count(ctx):
    """Every tile object count 0..33 (33 past the 32-entry capacity), so
    every lanes-per-object class of the tiled march runs: whole-group
    samples (1 object), 4/2-lane groups (2-4), one lane per object over 1-4
    passes, and the one-object last pass of 9, 17 and 25 objects. The busy
    counts go to the tiles the scene's objects cover, their culled objects
    last in the list (so the last object is a real one) after objects that
    miss the tile; some lists repeat their first object as their last (ties:
    the earlier entry wins) or name an object index past the scene's
    objects. Bit-exact against the FP32 twin given the same lists."""
    sc = scenes.spheres_scene(seed=4, n=40, width=144, height=80, half_cube=4.0, cam_dist=12.0)
    _say("case: host tile lists, every count 0..33, 144x80, 40 spheres")
    gs, tc, cf, ext = _upload(ctx, sc)
    ntx, nty = 9, 5
    idx, cc = O.tile_cull(sc.cam, sc.su, sc.objects, sc.width, sc.height)
    cull = idx.reshape(-1, 32)
    first = [9, 17, 25, 33, 32, 1, 2, 3, 4, 5, 8, 16, 24, 10, 18, 26]
    by_count = first + [c for c in range(34) if c not in first] + list(range(34))
    lists = np.zeros((ntx * nty, 32), np.uint32)
    counts = np.zeros(ntx * nty, np.uint32)
    for rank, t in enumerate(np.argsort(-cc.astype(np.int64), kind="stable")):
        c = by_count[rank]
        k = min(c, 32)
        real = [int(x) for x in cull[t, :min(int(cc[t]), 32)]]
        fill = [j for j in range(40) if j not in real]
        objs = ((fill[:max(0, k - len(real))] + real)[-k:] if k else []) + fill + list(range(40))
        objs = objs[:32]
        if k > 1 and rank % 3 == 0:
            objs[k - 1] = objs[0]
        if k > 2 and rank % 5 == 0:
            objs[k // 2] = 40 + int(t)
        lists[t] = objs
        counts[t] = c
    assert sorted(set(counts.tolist())) == list(range(34))
    tc.set_lists(lists.reshape(-1), counts)
    gb = M.GBuffer(ctx, sc.width, sc.height, debug=True)
    M.RayMarchPass(ctx).dispatch(gs, gb, tc, cf)
    cells, u = cf.read()
    inp = O.RefInputs(sc.pool, sc.maps, sc.objects, sc.cam, sc.su, ext, sc.width, sc.height,
                      tile_idx=np.ascontiguousarray(lists.reshape(-1)), tile_counts=counts,
                      coarse=cells, coarse_u=u)
    res = O.ref_trace(inp, prec=32)
    py, px = np.divmod(np.arange(sc.width * sc.height), sc.width)
    hits = np.bincount((py // 16) * ntx + px // 16, weights=res["hit"].astype(np.float64), minlength=ntx * nty)
    for c in (9, 17, 25, 32, 33):
        assert hits[counts == c].sum() > 0, c
    _compare(gb.read(), res, sc.width, sc.height)


def test_host_coarse_fields(ctx):
    """Coarse fields handed over with rkfm_set_coarse_field, including ones
    that are not distance fields (any field gives the FP32 twin's result bit
    for bit): the GPU-built field; that field with noise of up to +-0.6 and
    spikes of 40 on random cells (different hits); the field scaled by 0.25;
    a field with an infinite cell; and a negative inv_voxel_size in the
    uniforms (every point reads as outside the field)."""
    sc = scenes.spheres_scene(seed=6, n=24, width=160, height=96, half_cube=4.0, cam_dist=11.0)
    _say("case: host coarse fields 160x96, 24 spheres")
    gs, tc, cf, ext = _upload(ctx, sc)
    idx, counts = tc.read()
    base, u = cf.read()
    rng = np.random.default_rng(11)
    f = base.view(np.float16).astype(np.float32)
    noisy = f + rng.uniform(-0.6, 0.6, f.size).astype(np.float32)
    noisy[rng.random(f.size) < 0.01] = 40.0
    inf = f.copy()
    inf[f.size // 2] = np.inf
    u_neg = u.copy()
    u_neg["inv_voxel_size"] = -u["inv_voxel_size"]
    cases = (("built", base, u), ("noise and spikes", np.abs(noisy).astype(np.float16).view(np.uint16), u),
             ("scaled 0.25", (f * 0.25).astype(np.float16).view(np.uint16), u),
             ("infinite cell", inf.astype(np.float16).view(np.uint16), u),
             ("negative inv_voxel_size", base, u_neg))
    for name, cells, uu in cases:
        _say(f"case: host coarse field: {name}")
        cells = np.ascontiguousarray(cells, np.uint16)
        cf.upload_cells(cells, uu)
        gb = M.GBuffer(ctx, sc.width, sc.height, debug=True)
        M.RayMarchPass(ctx).dispatch(gs, gb, tc, cf)
        inp = O.RefInputs(sc.pool, sc.maps, sc.objects, sc.cam, sc.su, ext, sc.width, sc.height,
                          tile_idx=idx, tile_counts=counts, coarse=cells, coarse_u=uu)
        res = O.ref_trace(inp, prec=32)
        assert res["hit"].sum() > 100
        _compare(gb.read(), res, sc.width, sc.height)


def test_drop_in_1080p_parity(ctx):
    """The tiled drop-in at 1920x1080 (8,160 tiles in their object-count
    launch order, each tile's 32 workgroups on one XCD, a partial last tile
    row): every pixel bit-exact against the FP32 twin."""
    sc = scenes.spheres_scene(seed=0, width=1920, height=1080)
    _say("case: drop-in 1920x1080, 64 spheres")
    gs, tc, cf, ext = _upload(ctx, sc, L.STRATEGY_TILED)
    gb = M.GBuffer(ctx, sc.width, sc.height, debug=True)
    M.RayMarchPass(ctx).dispatch(gs, gb, tc, cf)
    idx, counts = tc.read()
    assert counts.size == 120 * 68 and counts.max() >= 7
    cells, u = cf.read()
    inp = O.RefInputs(sc.pool, sc.maps, sc.objects, sc.cam, sc.su, ext, sc.width, sc.height,
                      tile_idx=idx, tile_counts=counts, coarse=cells, coarse_u=u)
    res32, _ = O.ref_trace_parallel(inp)
    assert res32["hit"].sum() > 100000
    _compare(gb.read(), res32, sc.width, sc.height)

"""Batched launch: every batch (views, shadow-ray batches, point queries, over
two resident scenes) equals the same batch run alone, bit for bit, under every
launch shape and schedule, and across HIP graph replays interleaved with
other work."""
import numpy as np
import pytest

from oracle import oracle_py as O
from rkifield_amd import layouts as L
from rkifield_amd import march as M
from rkifield_amd import scenes

pytestmark = pytest.mark.gpu

TUNINGS = [(L.SCHEDULE_GRID, 8, 8, 0), (L.SCHEDULE_GRID, 32, 2, 0), (L.SCHEDULE_GRID, 4, 16, 0),
           (L.SCHEDULE_PERSISTENT, 16, 4, 16), (L.SCHEDULE_PERSISTENT, 8, 8, 4),
           (L.SCHEDULE_GRID_LOWLAT, 8, 8, 0), (L.SCHEDULE_GRID_LOWLAT, 16, 4, 0)]


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def two_scenes(ctx):
    out = []
    for seed in (11, 12):
        prims = scenes.primitive_field(seed=seed, n=40, half_cube=3.0, size_range=(0.3, 1.5), group=4)
        lo, hi = scenes.prim_bounds(prims, 0.5)
        out.append(M.SparseScene.build(ctx, prims, 0.05, lo, hi, smin_k=0.1, mat_seed=scenes.material_seed(seed)))
    yield out
    for s in out:
        s.destroy()


def _ext(shadows, cull=False):
    light = np.array([0.5, 1.0, 0.3])
    return L.scene_ext(hit_eps_mode=L.EPS_RELATIVE, shadow_enable=1 if shadows else 0,
                       light_dir=light / np.linalg.norm(light), march_flags=L.MARCH_TILE_CULL if cull else 0)


def _make(ctx, two_scenes):
    rng = np.random.default_rng(5)
    b = M.Batch(ctx)
    views = []
    # views 0 and 3 with per-tile ray intervals: each must march exactly as it does alone
    for k, (w, h, sc, sh, cull) in enumerate([(72, 40, 0, True, True), (33, 57, 1, False, False),
                                              (64, 64, 1, True, False), (50, 30, 1, True, True),
                                              (96, 64, 0, True, True)]):
        eye = rng.uniform(4, 7, 3) * rng.choice([-1, 1], 3)
        # the last view looks past the scene from afar: part of its frame
        # meets no occupied cell (empty interval tiles beside busy ones)
        cam = M.camera_look_at(eye, (0, 0, 0), 60.0, w, h) if k < 4 else \
            M.camera_look_at((14.0, 5.0, 9.0), (5.0, 2.5, 0.0), 60.0, w, h)
        su = L.scene_uniforms(num_objects=1)
        gb = M.GBuffer(ctx, w, h, debug=True, shadow=True)
        b.add_view(two_scenes[sc], cam, su, _ext(sh, cull), gb)
        views.append((two_scenes[sc], cam, su, _ext(sh, cull), w, h, gb))
    shadows = []
    for sc, n in ((0, 1000), (1, 77)):
        o = rng.uniform(-4, 4, (n, 3)).astype(np.float32)
        od, vd = ctx.to_device(o), ctx.array((n,), np.float32)
        b.add_shadow(two_scenes[sc], od, vd, n, _ext(True))
        shadows.append((two_scenes[sc], o, vd))
    points = []
    for sc, n in ((1, 3000), (0, 65)):
        x = rng.uniform(-5, 5, (n, 3)).astype(np.float32)
        xd, rd = ctx.to_device(x), ctx.array((n,), L.POINT_RESULT)
        b.add_points(two_scenes[sc], xd, rd, n)
        points.append((two_scenes[sc], x, xd, rd))
    for sc, n in ((0, 2000), (1, 33)):
        casts = _casts(rng, n)
        cd, rd = ctx.to_device(casts), ctx.array((n,), L.CAST_RESULT)
        b.add_spheres(two_scenes[sc], cd, rd, n)
        points.append((two_scenes[sc], casts, cd, rd))
    return b, views, shadows, points


def _casts(rng, n):
    c = np.zeros(n, L.SPHERE_CAST)
    c["origin"] = rng.uniform(-5, 5, (n, 3))
    d = rng.normal(size=(n, 3))
    c["dir"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    c["radius"] = rng.uniform(0.02, 0.6, n)
    c["max_t"] = 10.0
    return c


def _alone(ctx, views, points):
    ref_views = []
    for sc, cam, su, ext, w, h, _ in views:
        gb = M.GBuffer(ctx, w, h, debug=True, shadow=True)
        sc.march(cam, su, ext, gb, variant=0)
        ref_views.append(gb.read())
    ref_points = []
    for sc, x, xd, _ in points:
        if x.dtype == L.SPHERE_CAST:
            out = ctx.array((x.shape[0],), L.CAST_RESULT)
            sc.sphere_cast(xd, out, x.shape[0])
        else:
            out = ctx.array((x.shape[0],), L.POINT_RESULT)
            sc.query(xd, out, x.shape[0])
        ref_points.append(out.download())
    return ref_views, ref_points


def _check(views, shadows, points, ref_views, ref_sh, ref_points):
    for (sc, cam, su, ext, w, h, gb), ref in zip(views, ref_views):
        got = gb.read()
        for k in ("position", "normal", "material", "motion", "debug", "shadow"):
            assert np.array_equal(got[k].view(np.uint8), ref[k].view(np.uint8)), k
    for (sc, o, vd), ref in zip(shadows, ref_sh):
        assert np.array_equal(vd.download().view(np.uint32), ref.view(np.uint32))
    for (sc, x, xd, rd), ref in zip(points, ref_points):
        assert np.array_equal(rd.download().view(np.uint8), ref.view(np.uint8))


def _clear(views, shadows, points):
    for v in views:
        gb = v[-1]
        for plane in (gb.position, gb.normal, gb.material, gb.motion, gb.debug, gb.shadow):
            plane.fill(0xAB)
    for s in shadows:
        s[2].fill(0xAB)
    for p in points:
        p[3].fill(0xAB)


def _mixed_packets(t_lo, w, h, tw, th):
    """Number of tw x th pixel packets that hold pixels of both an empty
    8x8 interval tile (t_lo = +inf) and a non-empty one."""
    tx = (w + 7) // 8
    empty = np.isinf(t_lo.reshape(-1, tx))
    px, py = np.meshgrid(np.arange(w), np.arange(h))
    e = empty[py // 8, px // 8]
    n = 0
    for y0 in range(0, h, th):
        for x0 in range(0, w, tw):
            blk = e[y0:y0 + th, x0:x0 + tw]
            n += int(blk.any() and not blk.all())
    return n


def test_culled_views_mix_empty_and_busy_tiles(ctx, two_scenes):
    """The culled views' frames hold empty 8x8 tiles (t_lo = +inf) next to
    non-empty ones inside one packet for the 32x2, 4x16 and 16x4 shapes, so
    the batched parity below covers lanes that leave early beside lanes that
    march."""
    b, views, shadows, points = _make(ctx, two_scenes)
    mixed = {(32, 2): 0, (4, 16): 0, (16, 4): 0}
    for sc, cam, su, ext, w, h, _ in views:
        if not ext["march_flags"]:
            continue
        iv = sc.tile_intervals(cam, su, w, h).reshape(-1, 2)
        assert np.isfinite(iv[:, 0]).any()
        for tw, th in mixed:
            mixed[(tw, th)] += _mixed_packets(iv[:, 0], w, h, tw, th)
    assert all(v > 0 for v in mixed.values()), mixed


def test_batched_equals_alone(ctx, two_scenes):
    b, views, shadows, points = _make(ctx, two_scenes)
    ref_views, ref_points = _alone(ctx, views, points)
    ref_sh = []
    for sc, o, vd in shadows:
        l2, l1, dist, mat = sc.download()
        si = O.SparseInputs(sc.info["origin"], float(sc.info["voxel_size"]), sc.info["cells"], l2, l1, dist, mat)
        ref_sh.append(O.sparse_shadow(si, o, _ext(True)["light_dir"]))
    assert sum(int((r["position"][:, 3] != L.MAX_FLOAT).sum()) for r in ref_views) > 200
    for t in TUNINGS:
        _say(f"case: batch tuning {t}")
        _clear(views, shadows, points)
        b.run(tuning=t)
        ctx.sync()
        _check(views, shadows, points, ref_views, ref_sh, ref_points)


def test_graph_replays_and_autotune(ctx, two_scenes):
    b, views, shadows, points = _make(ctx, two_scenes)
    ref_views, ref_points = _alone(ctx, views, points)
    b.run(tuning=TUNINGS[0])
    ctx.sync()
    ref_sh = [s[2].download() for s in shadows]
    for t in (TUNINGS[0], TUNINGS[3], TUNINGS[5]):
        _say(f"case: graph tuning {t}")
        g = b.graph(tuning=t)
        for rep in range(3):
            _clear(views, shadows, points)
            ctx.sync()
            g.launch()
            ctx.sync()
            _check(views, shadows, points, ref_views, ref_sh, ref_points)
            # unrelated eager work between replays
            views[0][0].march(views[0][1], views[0][2], views[0][3], M.GBuffer(ctx, 16, 16), variant=0)
            ctx.sync()
        g.destroy()
    best, ms = b.autotune(reps=2)
    assert best["tile_w"] * best["tile_h"] == 64 and ms > 0
    _clear(views, shadows, points)
    b.run()
    ctx.sync()
    _check(views, shadows, points, ref_views, ref_sh, ref_points)


def test_batch_errors(ctx, two_scenes):
    b = M.Batch(ctx)
    with pytest.raises(M.RkfMarchError):
        b.run()
    gb = M.GBuffer(ctx, 8, 8)
    b.add_view(two_scenes[0], M.camera_look_at((5, 5, 5), (0, 0, 0), 60.0, 8, 8), L.scene_uniforms(), _ext(False), gb)
    with pytest.raises(M.RkfMarchError):
        b.run(tuning=(L.SCHEDULE_GRID, 8, 4, 0))  # 32 lanes: invalid packet
    with pytest.raises(M.RkfMarchError):
        b.run(tuning=(3, 8, 8, 0))  # no such schedule
    # tile culling is a view option only
    b2 = M.Batch(ctx)
    vd = ctx.array((4,), np.float32)
    b2.add_shadow(two_scenes[0], ctx.to_device(np.zeros((4, 3), np.float32)), vd, 4, _ext(True, cull=True))
    with pytest.raises(M.RkfMarchError):
        b2.run()

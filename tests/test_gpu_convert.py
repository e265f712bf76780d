"""Reference scene -> resident sparse scene (rkfm_sparse_from_objects): the
converter's index, distance, material and aux planes equal the oracle's
restatement byte for byte (mixed analytic / rotated / scaled / geometry-first
voxelised objects, and the 64 voxelised spheres); the sparse march over the
converted scene is bit-exact against the sparse oracle; and against the
drop-in march of the original scene (FP64 oracle, reference semantics) it
finds the same hits, within one voxel in depth."""
import numpy as np
import pytest

from oracle import oracle_py as O
from parity_util import compare_sparse, say
from rkifield_amd import layouts as L
from rkifield_amd import march as M
from rkifield_amd import scenes

pytestmark = pytest.mark.gpu


def _upload(ctx, sc):
    gs = M.GpuScene(ctx)
    gs.upload_brick_pool(sc.pool)
    gs.upload_brick_maps(sc.maps)
    gs.upload_objects(sc.objects)
    gs.update_camera(sc.cam)
    gs.update_scene(sc.su, sc.ext)
    return gs


@pytest.mark.parametrize("which", ["objects", "spheres"])
def test_from_objects_matches_oracle(ctx, which):
    sc = scenes.objects_scene() if which == "objects" else scenes.spheres_scene(seed=0, width=320, height=180)
    b = 4.5 if which == "objects" else 9.0
    lo, hi = (-b, -b, -b), (b, b, b)
    say(f"case: from_objects {which}, vs {sc.voxel_size}, bounds +-{b}")
    _upload(ctx, sc)
    sp = M.SparseScene.from_objects(ctx, sc.voxel_size, lo, hi)
    try:
        l2, l1, dist, mat = sp.download()
        aux = sp.download_aux()
        inp = O.RefInputs(sc.pool, sc.maps, sc.objects, sc.cam, sc.su, sc.ext, sc.width, sc.height)
        cells, ol2, ol1, odist, omat, oaux = O.sparse_from_objects(inp, sc.voxel_size, lo, hi)
        assert list(sp.info["cells"]) == list(cells)
        assert np.array_equal(l2, ol2) and np.array_equal(l1, ol1)
        assert np.array_equal(dist, odist) and np.array_equal(mat, omat)
        if which == "objects":
            assert aux is not None and np.array_equal(aux, oaux) and (aux != 0).sum() > 1000
        else:
            assert aux is None and not oaux.any()
        # the sparse march over the converted scene
        si = O.SparseInputs(sp.info["origin"], float(sp.info["voxel_size"]), sp.info["cells"], l2, l1, dist, mat,
                            aux=aux)
        light = np.array([0.5, 1.0, 0.3])
        ext = L.scene_ext(hit_eps_mode=L.EPS_RELATIVE, shadow_enable=1, light_dir=light / np.linalg.norm(light),
                          march_flags=L.MARCH_TILE_CULL)
        cam = sc.cam.copy()
        gb = M.GBuffer(ctx, sc.width, sc.height, debug=True, shadow=True)
        sp.march(cam, sc.su, ext, gb)
        gpu = gb.read()
        res = O.sparse_trace(si, cam, sc.su, ext, sc.width, sc.height, prec=32)
        hits = compare_sparse(gpu, res)
        assert hits > 1000
        if which == "spheres":
            # against the drop-in's reference semantics on the original scene
            ext_b = sc.ext.copy()
            ext_b["strategy"] = L.STRATEGY_BRUTE
            ref = O.ref_trace(O.RefInputs(sc.pool, sc.maps, sc.objects, sc.cam, sc.su, ext_b, sc.width, sc.height),
                              prec=64)
            ext64 = sc.ext.copy()
            r64 = O.sparse_trace(si, cam, sc.su, ext64, sc.width, sc.height, prec=64)
            h1, h2 = r64["hit"] == 1, ref["hit"] == 1
            both = h1 & h2
            dt = np.abs(r64["t"][both] - ref["t"][both])
            ang = np.arccos(np.clip(np.sum(r64["normal"][both] * ref["normal"][both], axis=1), -1, 1))
            say(f"converted vs original: hit agreement {(h1 == h2).mean():.5f}, depth |dt| max {dt.max():.4f} m, "
                f"normal angle p99 {np.percentile(ang, 99):.4f} rad")
            assert (h1 == h2).mean() > 0.999
            assert dt.max() < sc.voxel_size
            assert np.percentile(ang, 99) < 0.05
    finally:
        sp.destroy()

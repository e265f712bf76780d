"""Parity of the terrain stress workload at its own size: 1 km^2 fBm terrain
+ 20k smooth-min rocks (24 M bricks at 5 cm). One frame of 32 CSG sphere
edits (cpu_apply.rs:28-188 semantics) leaves every distance and material
byte of the pool and the modified-brick list equal to the oracle's; then the
3840x2160 frame (+ one shadow ray per hit, tile intervals) is bit-exact
against the FP32 twin over the edited scene, with FP64 depth and normal
errors over the 24 rows with most hits."""
import time

import numpy as np
import pytest

from oracle import oracle_py as O
from parity_util import compare_sparse, fp64_rows, say, sparse_inputs
from rkifield_amd import march as M
from rkifield_amd import scenes

pytestmark = pytest.mark.gpu


def test_terrain_edits_and_4k_frame(ctx):
    prims, lo, hi = scenes.terrain_scene(seed=4)
    t0 = time.time()
    sc = M.SparseScene.build(ctx, prims, 0.05, lo, hi, smin_k=0.1, mat_seed=scenes.material_seed(4))
    try:
        si = sparse_inputs(sc)
        say(f"case: terrain {int(sc.info['n_bricks'])} bricks, oracle copy in {time.time() - t0:.1f} s")
        ops = scenes.terrain_edits(prims[0], scenes.Pcg32(44))
        mod = sc.edit(ops)
        _, _, omod = O.sparse_edit(si, ops, in_place=True)
        assert np.array_equal(mod, omod) and mod.size > 1000
        _, _, dist, mat = sc.download()
        assert np.array_equal(dist, si.dist), f"{np.sum(dist != si.dist)} distances differ"
        assert np.array_equal(mat, si.mat), f"{np.sum(mat != si.mat)} materials differ"
        del dist, mat
        say(f"terrain: 32 edits, {mod.size} modified bricks, every distance/material byte equal")
        W, H = 3840, 2160
        cam, su, ext = scenes.terrain_view(prims[0], W, H)
        gb = M.GBuffer(ctx, W, H, debug=True, shadow=True)
        sc.march(cam, su, ext, gb)
        gpu = gb.read()
        res, _ = O.sparse_trace_parallel(si, cam, su, ext, W, H)
        hits = compare_sparse(gpu, res)
        same, rel, ang = fp64_rows(si, cam, su, ext, W, H, gpu)
        say(f"terrain 4K: {hits} hits bit-exact; FP64 over 24 rows: same-path {same:.5f}, max rel depth "
            f"{rel.max():.3e}, max normal angle {ang.max():.3e} rad")
        assert hits > 1000000 and same > 0.99
        assert rel.max() <= 1e-4 and ang.max() <= 1e-3
    finally:
        sc.destroy()

"""rkfm_sparse_march_noalloc (the allocation-free sparse march, scratch in
the library's fixed per-device array) against rkfm_sparse_march through the
C ABI: every plane bit for bit, on two calls in a row, the second with a
smaller frame over a smaller scene, culled and unculled, both variants; and
the size limits are refused before anything runs."""
import numpy as np
import pytest

from rkifield_amd import layouts as L
from rkifield_amd import march as M
from rkifield_amd import scenes

pytestmark = pytest.mark.gpu

PLANES = ("position", "normal", "material", "motion", "debug", "shadow")


def _say(msg):
    print(msg, flush=True)


def _scene(ctx, seed, n, half):
    prims = scenes.primitive_field(seed=seed, n=n, half_cube=half, size_range=(0.3, 1.5), group=4)
    lo, hi = scenes.prim_bounds(prims, 0.5)
    return M.SparseScene.build(ctx, prims, 0.05, lo, hi, smin_k=0.1, mat_seed=scenes.material_seed(seed))


def _ext(cull):
    light = np.array([0.5, 1.0, 0.3])
    return L.scene_ext(hit_eps_mode=L.EPS_RELATIVE, shadow_enable=1, light_dir=light / np.linalg.norm(light),
                       march_flags=L.MARCH_TILE_CULL if cull else 0)


def test_noalloc_march_equals_march(ctx):
    big, small = _scene(ctx, 41, 48, 3.0), _scene(ctx, 42, 12, 1.5)
    try:
        su = L.scene_uniforms(num_objects=1)
        calls = [(big, 104, 72, (6.0, 3.0, 7.0), True, 0), (small, 40, 24, (3.0, 2.0, 3.5), True, 0),
                 (big, 64, 48, (-5.0, 2.0, 6.0), False, 0), (small, 33, 17, (2.5, -1.5, 3.0), True, 1)]
        hits = 0
        for sc, w, h, eye, cull, variant in calls:
            _say(f"case: noalloc {w}x{h} cull {cull} variant {variant}")
            cam = M.camera_look_at(eye, (0, 0, 0), 60.0, w, h)
            want = M.GBuffer(ctx, w, h, debug=True, shadow=True)
            sc.march(cam, su, _ext(cull), want, variant=variant)
            got = M.GBuffer(ctx, w, h, debug=True, shadow=True)
            for plane in (got.position, got.normal, got.material, got.motion, got.debug, got.shadow):
                plane.fill(0xAB)
            sc.march_noalloc(cam, su, _ext(cull), got, variant=variant)
            a, b = want.read(), got.read()
            for k in PLANES:
                assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), k
            hits += int((a["position"][:, 3] != L.MAX_FLOAT).sum())
        assert hits > 1000
        # limits: refused before anything is launched
        cam = M.camera_look_at((6.0, 3.0, 7.0), (0, 0, 0), 60.0, 4104, 4104)
        gb = M.GBuffer(ctx, 4104, 4104)
        with pytest.raises(M.RkfMarchError) as e:
            big.march_noalloc(cam, su, _ext(True), gb)
        assert e.value.code == M.ERR["OUT_OF_RANGE"]
        with pytest.raises(M.RkfMarchError) as e:
            big.march_noalloc(cam, su, _ext(True), gb, variant=2)
        assert e.value.code == M.ERR["INVALID_ARG"]
    finally:
        big.destroy()
        small.destroy()

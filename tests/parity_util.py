"""Shared helpers of the GPU parity tests (not a test module): upload a
reference-layout scene, compare the G-buffer with the oracle's FP32 twin bit
for bit, and measure depth / normal errors against the FP64 oracle."""
import numpy as np

from oracle import oracle_py as O
from rkifield_amd import layouts as L
from rkifield_amd import march as M


def say(msg):
    print(msg, flush=True)


def upload_ref(ctx, sc, strategy=L.STRATEGY_AUTO, tiles=True, coarse=True):
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


def ref_inputs(sc, ext, tc=None, cf=None, tiles=True, coarse=True):
    kw = dict(bvh=M.build_bvh(sc.objects))
    if tiles and tc is not None:
        idx, counts = tc.read()
        kw.update(tile_idx=idx, tile_counts=counts)
    if coarse and cf is not None:
        cells, u = cf.read()
        kw.update(coarse=cells, coarse_u=u)
    return O.RefInputs(sc.pool, sc.maps, sc.objects, sc.cam, sc.su, ext, sc.width, sc.height, **kw)


def compare_ref(gpu, res):
    """Bit-exact comparison of the drop-in kernel's G-buffer and debug plane
    with the FP32 twin."""
    pos, nrm, mat, mot = O.pack_gbuffer(res)
    dbg = gpu["debug"]
    hit = res["hit"] == 1
    ghit = gpu["position"][:, 3] != L.MAX_FLOAT
    assert np.array_equal(ghit, hit), f"hit mask differs on {np.sum(ghit != hit)} rays"
    assert np.array_equal(gpu["material"], mat), f"material differs on {np.sum(gpu['material'] != mat)} rays"
    for f in ("brick_slot", "voxel_index", "cell_index", "steps", "evals"):
        assert np.array_equal(dbg[f], res[f]), f
    assert np.array_equal(gpu["position"].view(np.uint32), pos.view(np.uint32))
    assert np.array_equal(gpu["normal"], nrm)
    assert np.array_equal(gpu["motion"].view(np.uint32), mot.view(np.uint32))
    dn = np.asarray(res["normal"], np.float32)
    assert np.array_equal(dbg["normal"][hit].view(np.uint32), dn[hit].view(np.uint32))


def depth_normal_errors(t_gpu, n_gpu, res64, res32):
    """Over rays whose discrete path (hit, step count and the hash of every
    branch decision of the march) is the same in FP32 and FP64: (same-path
    share, relative depth errors, normal angles). A ray whose FP32 and FP64
    decisions differ (a sample on either side of a threshold or of a free
    region's face) is counted outside the same-path share."""
    same = (res64["hit"] == res32["hit"]) & (res64["steps"] == res32["steps"]) & (res64["path"] == res32["path"])
    hit = same & (res32["hit"] == 1) & (res64["t"] > 0)
    t = np.asarray(t_gpu, np.float64)
    rel = np.abs(t[hit] - res64["t"][hit]) / res64["t"][hit]
    n = np.asarray(n_gpu, np.float64)
    cosang = np.clip(np.sum(n[hit] * res64["normal"][hit], axis=1), -1.0, 1.0)
    return same.mean(), rel, np.arccos(cosang)


SPARSE_DEBUG = ("brick_slot", "voxel_index", "cell_index", "steps", "evals", "shadow_steps", "shadow_evals")


def compare_sparse(gpu, res, shadow=True):
    """Bit-exact comparison of a sparse-scene G-buffer (+ debug, shadow
    planes) with the FP32 twin; returns the number of hits."""
    pos, nrm, mat, mot = O.pack_gbuffer(res)
    hit = res["hit"] == 1
    ghit = gpu["position"][:, 3] != L.MAX_FLOAT
    assert np.array_equal(ghit, hit), f"hit mask differs on {np.sum(ghit != hit)} rays"
    assert np.array_equal(gpu["material"], mat), f"material differs on {np.sum(gpu['material'] != mat)} rays"
    if "debug" in gpu:
        for f in SPARSE_DEBUG:
            assert np.array_equal(gpu["debug"][f], res[f]), f
    assert np.array_equal(gpu["position"].view(np.uint32), pos.view(np.uint32))
    assert np.array_equal(gpu["normal"], nrm)
    assert np.array_equal(gpu["motion"].view(np.uint32), mot.view(np.uint32))
    if shadow:
        assert np.array_equal(gpu["shadow"].view(np.uint32), res["shadow"].astype(np.float32).view(np.uint32))
    return int(hit.sum())


def sparse_inputs(sc, aux=None):
    l2, l1, dist, mat = sc.download()
    i = sc.info
    return O.SparseInputs(i["origin"], float(i["voxel_size"]), i["cells"], l2, l1, dist, mat, aux=aux)


def fp64_rows(si, cam, su, ext, w, h, gpu, n_rows=24):
    """FP64 oracle over the n_rows rows with the most GPU hits: (same-path
    share, relative depth errors, normal angles) of those rows."""
    hits = (gpu["position"][:, 3] != L.MAX_FLOAT).reshape(h, w).sum(axis=1)
    rows = np.sort(np.argsort(-hits, kind="stable")[:n_rows])
    sames, rels, angs = [], [], []
    for r in rows:
        p0, p1 = int(r) * w, (int(r) + 1) * w
        r64 = O.sparse_trace(si, cam, su, ext, w, h, prec=64, p0=p0, p1=p1)
        r32 = O.sparse_trace(si, cam, su, ext, w, h, prec=32, p0=p0, p1=p1)
        s, rel, ang = depth_normal_errors(gpu["position"][p0:p1, 3], gpu["debug"]["normal"][p0:p1], r64, r32)
        sames.append(s)
        rels.append(rel)
        angs.append(ang)
    return float(np.mean(sames)), np.concatenate(rels), np.concatenate(angs)

"""Benchmark of the SDF ray-march hot path on one MI355X.

Default workload (the metric's 1080p configuration): 4096 primitives
(sphere/box/capsule, size U[0.1,4] m full extent, random rotation, smooth-min
k=0.1 in groups of 8) in a 256 m cube, voxelised at 0.02 m into a resident
two-level sparse scene; 1920x1080 primary rays plus one soft-shadow ray per hit
toward normalize(0.5, 1, 0.3); 128 near steps, hit eps 1e-3 t, far 100 m.

A step is one launch of the march over the frame (primary + shadow rays),
inputs resident in HBM. Timed with HIP events on the launch stream. The
algorithmic bytes of a launch come from the evaluation counts the kernel
reports in its debug plane (equal to the oracle's, checked by the parity
tests): 36 B per SDF evaluation (8 corner words of 4 B + one 4 B index
entry), 6 * 36 + 20 B per hit (normal, material/blend voxel, map), 44 B
G-buffer write per pixel, 4 B per shadow-ray result.

--workload config0 runs the drop-in kernel on the reference layouts instead
(64 voxelised spheres, 640x360); config2 the batched graph-captured launch;
config3 the served request stream; config4 the 4K terrain frame with 32 CSG
edits per frame.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def build_config1(ctx, M, L, scenes, width, height):
    prims = scenes.primitive_field(seed=1, n=4096, half_cube=128.0, size_range=(0.1, 4.0), group=8)
    lo, hi = scenes.prim_bounds(prims, 1.0)
    t0 = time.time()
    sc = M.SparseScene.build(ctx, prims, 0.02, lo, hi, smin_k=0.1, mat_seed=scenes.material_seed(1))
    ctx.sync()
    log(f"config1 scene: {int(sc.info['n_bricks'])} bricks, {int(sc.info['n_blocks'])} L1 blocks, "
        f"cells {list(sc.info['cells'])}, {sc.info['device_bytes'] / 2**30:.2f} GiB, built in {time.time() - t0:.1f} s")
    cam, su, ext = scenes.config1_view(width, height, cull=False)
    return sc, cam, su, ext


def algorithmic_bytes(dbg, hit):
    evals = dbg["evals"].astype(np.int64)
    sh = dbg["shadow_evals"].astype(np.int64)
    b = 36 * evals.sum() + 44 * evals.size + int(hit.sum()) * (6 * 36 + 20)
    b += 36 * sh.sum() + 4 * int(hit.sum())
    return int(b)


def accuracy_sparse(O, si, cam, su, ext, width, height, gpu, res, r0, rows, n64=8):
    """GPU frame against the oracle: every field of rows r0..r0+rows-1 against
    the FP32 twin (count of differing pixels), and the FP64 oracle over the
    n64 rows with most hits (rays whose discrete path is the same in FP32 and
    FP64): max relative depth error, max normal angle."""
    import numpy as np
    from rkifield_amd import layouts as L
    p0, p1 = r0 * width, (r0 + rows) * width
    pos, nrm, mat, mot = O.pack_gbuffer(res)
    g = {k: v[p0:p1] for k, v in gpu.items()}
    bad = ~((g["position"].view(np.uint32) == pos.view(np.uint32)).all(axis=1)
            & (g["normal"] == nrm).all(axis=1) & (g["material"] == mat)
            & (g["motion"].view(np.uint32) == mot.view(np.uint32)).all(axis=1)
            & (g["shadow"].view(np.uint32) == res["shadow"].astype(np.float32).view(np.uint32)))
    for f in ("brick_slot", "voxel_index", "cell_index", "steps", "evals", "shadow_steps", "shadow_evals"):
        bad |= g["debug"][f] != res[f]
    hits = (gpu["position"][:, 3] != L.MAX_FLOAT).reshape(height, width).sum(axis=1)
    rows64 = np.sort(np.argsort(-hits, kind="stable")[:n64])
    rels, angs, same_n, tot = [], [], 0, 0
    for r in rows64:
        q0, q1 = int(r) * width, (int(r) + 1) * width
        r64 = O.sparse_trace(si, cam, su, ext, width, height, prec=64, p0=q0, p1=q1)
        r32 = O.sparse_trace(si, cam, su, ext, width, height, prec=32, p0=q0, p1=q1)
        same = (r64["hit"] == r32["hit"]) & (r64["steps"] == r32["steps"]) & (r64["path"] == r32["path"])
        same_n += int(same.sum())
        tot += same.size
        h = same & (r32["hit"] == 1)
        t = gpu["position"][q0:q1, 3].astype(np.float64)
        rels.append(np.abs(t[h] - r64["t"][h]) / r64["t"][h])
        n = gpu["debug"]["normal"][q0:q1].astype(np.float64)
        angs.append(np.arccos(np.clip(np.sum(n[h] * r64["normal"][h], axis=1), -1.0, 1.0)))
    rel, ang = np.concatenate(rels), np.concatenate(angs)
    return dict(twin_rows=f"{r0}..{r0 + rows - 1}", twin_pixels=int(p1 - p0), twin_mismatches=int(bad.sum()),
                fp64_rows=[int(r) for r in rows64], fp64_hits=int(rel.size), same_path=round(same_n / tot, 6),
                max_rel_depth=float(rel.max(initial=0.0)), share_depth_1e4=round(float((rel <= 1e-4).mean()), 6),
                max_normal_rad=float(ang.max(initial=0.0)),
                share_normal_1e3=round(float((ang <= 1e-3).mean()), 6))


def cpu_baseline_sparse(sc, cam, su, ext, width, height, seconds, gpu=None):
    from oracle import oracle_py as O
    l2, l1, dist, mat = sc.download()
    si = O.SparseInputs(sc.info["origin"], float(sc.info["voxel_size"]), sc.info["cells"], l2, l1, dist, mat)
    # calibrate on 4 rows around the centre, then size a band (or repeat the
    # whole frame) for ~`seconds` of CPU work
    mid = height // 2
    t0 = time.time()
    res, used = O.sparse_trace_parallel(si, cam, su, ext, width, height, mid * width, (mid + 4) * width)
    dt = max(time.time() - t0, 1e-3)
    rows = int(max(4, min(height, 4 * seconds / dt)))
    r0 = max(0, mid - rows // 2)
    reps, rays, t0 = 0, 0, time.time()
    while True:
        res, used = O.sparse_trace_parallel(si, cam, su, ext, width, height, r0 * width, (r0 + rows) * width)
        rays += res.size + int(res["hit"].sum())
        reps += 1
        dt = time.time() - t0
        if dt >= seconds or reps >= 200:
            break
    acc = accuracy_sparse(O, si, cam, su, ext, width, height, gpu, res, r0, rows) if gpu is not None else None
    return dict(value=rays / dt / 1e6, unit="Mrays/s", cores=int(used), kind="port",
                sample=f"the oracle's FP32 twin of this path (C, OpenMP dynamic/64): the reference's brick "
                       f"sampling and step rules with the same free-space DDA and tile intervals as the kernel, "
                       f"{reps} pass(es) over rows {r0}..{r0 + rows - 1} of the 1920x1080 frame: {res.size} "
                       f"primary + {int(res['hit'].sum())} shadow rays per pass, {dt:.1f} s; host {cpu_model()}"), acc


def cpu_baseline_ref(sc, ext, idx, counts, cells, u, seconds):
    from oracle import oracle_py as O
    inp = O.RefInputs(sc.pool, sc.maps, sc.objects, sc.cam, sc.su, ext, sc.width, sc.height,
                      tile_idx=idx, tile_counts=counts, coarse=cells, coarse_u=u)
    reps, rays, t0 = 0, 0, time.time()
    while True:
        res, used = O.ref_trace_parallel(inp)
        rays += res.size
        reps += 1
        dt = time.time() - t0
        if dt >= seconds or reps >= 500:
            break
    return dict(value=rays / dt / 1e6, unit="Mrays/s", cores=int(used), kind="port",
                sample=f"the oracle's FP32 twin of ray_march_tiled (C, OpenMP): the reference's control flow "
                       f"and sampling over the same tile lists and coarse field, {reps} passes over the full "
                       f"640x360 frame in {dt:.1f} s; host {cpu_model()}"), res


def accuracy_ref(sc, ext, idx, counts, cells, u, gpu, twin):
    """Drop-in frame against the oracle: pixels whose G-buffer or debug
    fields differ from the FP32 twin, and the FP64 oracle's depth and normal
    errors over the rays whose discrete path is the same in FP32 and FP64."""
    from oracle import oracle_py as O
    inp = O.RefInputs(sc.pool, sc.maps, sc.objects, sc.cam, sc.su, ext, sc.width, sc.height,
                      tile_idx=idx, tile_counts=counts, coarse=cells, coarse_u=u)
    pos, nrm, mat, mot = O.pack_gbuffer(twin)
    bad = ~((gpu["position"].view(np.uint32) == pos.view(np.uint32)).all(axis=1)
            & (gpu["normal"] == nrm).all(axis=1) & (gpu["material"] == mat)
            & (gpu["motion"].view(np.uint32) == mot.view(np.uint32)).all(axis=1))
    for f in ("brick_slot", "voxel_index", "cell_index", "steps", "evals"):
        bad |= gpu["debug"][f] != twin[f]
    r64 = O.ref_trace(inp, prec=64)
    same = (r64["hit"] == twin["hit"]) & (r64["steps"] == twin["steps"]) & (r64["path"] == twin["path"])
    h = same & (twin["hit"] == 1)
    t = gpu["position"][:, 3].astype(np.float64)
    rel = np.abs(t[h] - r64["t"][h]) / r64["t"][h]
    n = gpu["debug"]["normal"].astype(np.float64)
    ang = np.arccos(np.clip(np.sum(n[h] * r64["normal"][h], axis=1), -1.0, 1.0))
    return dict(twin_pixels=int(twin.size), twin_mismatches=int(bad.sum()), fp64_hits=int(h.sum()),
                same_path=round(float(same.mean()), 6), max_rel_depth=float(rel.max(initial=0.0)),
                share_depth_1e4=round(float((rel <= 1e-4).mean()), 6), max_normal_rad=float(ang.max(initial=0.0)),
                share_normal_1e3=round(float((ang <= 1e-3).mean()), 6))


def run_config2(args, ctx, M, L, scenes):
    """64 uneven batches over the config1 scene in one graph-captured launch
    (autotuned shape) vs the same batches launched one by one."""
    sc, _, su, ext = build_config1(ctx, M, L, scenes, 1920, 1080)
    ext_v = ext.copy()  # views: per-tile ray intervals, as in the 1080p frame
    ext_v["march_flags"] = 0 if args.no_cull else L.MARCH_TILE_CULL
    views, sh_orig, pts = scenes.mixed_batches(seed=2)
    b = M.Batch(ctx)
    singles = []
    for (w, h, pos, tgt) in views:
        cam = M.camera_look_at(pos, tgt, 60.0, w, h)
        gb = M.GBuffer(ctx, w, h, shadow=True)
        b.add_view(sc, cam, su, ext_v, gb)
        singles.append(("view", cam, gb))
    for o in sh_orig:
        od, vd = ctx.to_device(o), ctx.array((o.shape[0],), np.float32)
        b.add_shadow(sc, od, vd, o.shape[0], ext)
        sb = M.Batch(ctx)
        sb.add_shadow(sc, od, vd, o.shape[0], ext)
        singles.append(("shadow", sb, None))
    for x in pts:
        xd, rd = ctx.to_device(x), ctx.array((x.shape[0],), L.POINT_RESULT)
        b.add_points(sc, xd, rd, x.shape[0])
        singles.append(("points", (xd, rd, x.shape[0]), None))
    work = b.work_items()
    t0 = time.time()
    best, best_ms = b.autotune(reps=3)
    log(f"autotune: {time.time() - t0:.1f} s -> schedule {int(best['schedule'])} tile {int(best['tile_w'])}x"
        f"{int(best['tile_h'])} waves/CU {int(best['waves_per_cu'])} ({best_ms:.3f} ms)")
    g = b.graph(tuning=best)
    for _ in range(args.warmup):
        g.launch()
    ctx.sync()
    ctx.timer_begin()
    for _ in range(args.steps):
        g.launch()
    ms = ctx.timer_end() / args.steps
    ctx.sync()

    def one_by_one():
        for kind, a, bb in singles:
            if kind == "view":
                sc.march(a, su, ext_v, bb, variant=0)
            elif kind == "shadow":
                a.run(tuning=(L.SCHEDULE_GRID, 8, 8, 0))
            else:
                sc.query(a[0], a[1], a[2])
    one_by_one()
    ctx.sync()
    ctx.timer_begin()
    for _ in range(args.steps):
        one_by_one()
    ms_single = ctx.timer_end() / args.steps
    ctx.sync()
    log(f"config2: {work} rays+queries per launch; batched graph {ms:.3f} ms, one launch per batch {ms_single:.3f} ms")
    line = {
        "metric": "Mrays/s (rays + queries) of 64 uneven batches in one graph-captured launch",
        "value": round(work / (ms / 1e3) / 1e6, 3), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded)",
        "config": {"workload": "config2: 16 views W~U{256..1920} H~U{256..1080} (+ shadow ray per hit), "
                               "16 shadow batches and 32 point batches of U{1k..64k}, config1 scene",
                   "tuning": {k: int(best[k]) for k in ("schedule", "tile_w", "tile_h", "waves_per_cu")},
                   "tile_cull": not args.no_cull, "work_items": work,
                   "one_launch_per_batch_ms": round(ms_single, 4)},
    }
    print(json.dumps(line), flush=True)


def build_scene(ctx, M, scenes, seed):
    prims = scenes.primitive_field(seed=seed, n=4096, half_cube=128.0, size_range=(0.1, 4.0), group=8)
    lo, hi = scenes.prim_bounds(prims, 1.0)
    return M.SparseScene.build(ctx, prims, 0.02, lo, hi, smin_k=0.1, mat_seed=scenes.material_seed(seed))


def served_vs_plain(ctx, M, L, scenes, resident, n_requests):
    """Replays the served trace (n_requests, Poisson 2000 req/s over the
    resident scenes) against the batched server and against the per-request
    policy: once all at once to warm up, at its arrival times (latency), then
    all at once three times (saturation: the median wall time).
    Returns (per-policy results, work items, offered Mrays/s, trace)."""
    n_scenes = len(resident)
    trace = scenes.served_trace(seed=3, n=n_requests, rate=2000.0, n_scenes=n_scenes)
    pool_n, ring_n, cap = 64, 256, 64 * 1024
    rng = np.random.default_rng(3)
    pool = M.PinnedArray((pool_n, cap, 3), np.float32)
    pool.a[:] = rng.uniform(-128.0, 128.0, (pool_n, cap, 3))
    ring = M.PinnedArray((ring_n, cap), L.POINT_RESULT)
    _, su, ext = scenes.config1_view(cull=False)
    reqs = np.zeros(len(trace), L.REQUEST)
    arrivals = np.zeros(len(trace))
    work = 0
    for i, (t, kind, sc, p) in enumerate(trace):
        arrivals[i] = t
        if kind == "points":
            r = reqs[i]
            r["kind"], r["scene"], r["n"] = L.BATCH_POINTS, sc, p
            r["in"] = pool.ptr + (i % pool_n) * cap * 12
            r["out"] = ring.ptr + (i % ring_n) * cap * L.POINT_RESULT.itemsize
            work += p
        else:
            w, h, pos, tgt = p
            cam = M.camera_look_at(pos, tgt, 60.0, w, h)
            reqs[i] = M.Server.view_request(sc, cam, su, ext, w, h, keep_on_device=True)
            work += w * h
    offered = work / arrivals[-1] / 1e6
    log(f"trace: {len(trace)} requests over {arrivals[-1]:.2f} s, {work / len(trace) / 1e3:.1f} k rays+queries "
        f"per request, offered {offered:.1f} Mrays/s")
    # a 4 KiB copy_kernel launch before and after each timed replay marks the
    # phases in a kernel trace of this run
    mk = (ctx.array((4096,), np.uint8), ctx.array((4096,), np.uint8))

    def mark():
        ctx.copy(mk[1], mk[0], 4096)
        ctx.sync()

    results = {}
    for name, policy in (("served", L.SERVE_BATCHED), ("plain", L.SERVE_PER_REQUEST)):
        srv = M.Server(ctx, resident, memory_budget=4 << 30, max_requests=ring_n, policy=policy,
                       max_pending=len(trace) + 16)
        # warm-up: the whole trace at once, so every scene and request kind
        # has been served before the timed replays
        srv.replay(reqs, np.zeros(len(trace)))
        st0 = srv.stats()
        mark()
        lat, wall = srv.replay(reqs, arrivals)
        mark()
        st = srv.stats()
        ok = np.isfinite(lat)
        lat_ms = lat[ok] * 1e3
        worst = np.argsort(-np.where(ok, lat, 0.0))[:5]
        log(f"{name}: slowest requests " + ", ".join(
            f"{lat[i] * 1e3:.2f} ms ({trace[i][1]} at {arrivals[i] * 1e3:.1f} ms)" for i in worst))
        # saturation: the whole trace at once, median of three replays
        walls = []
        for _ in range(3):
            walls.append(srv.replay(reqs, np.zeros(len(trace)))[1])
            mark()
        burst_wall = float(np.median(walls))
        st2 = srv.stats()
        srv.close()
        r = dict(p50_ms=float(np.percentile(lat_ms, 50)), p99_ms=float(np.percentile(lat_ms, 99)),
                 mean_ms=float(lat_ms.mean()), max_ms=float(lat_ms.max()), failed=int((~ok).sum()),
                 mrays_s_at_2k=work / wall / 1e6, rounds=int(st["rounds"] - st0["rounds"]),
                 saturation_mrays_s=work / burst_wall / 1e6, saturation_s=burst_wall,
                 saturation_rounds=int(st2["rounds"] - st["rounds"]) // 3,
                 saturation_walls_s=[round(w, 4) for w in walls],
                 gpu_s=float(st2["gpu_seconds"] - st0["gpu_seconds"]),
                 arena_peak_mib=float(st2["arena_peak"]) / 2**20)
        for kind in ("points", "view512", "view1080"):
            sel = np.array([(k == "points") if kind == "points" else (k == "view" and v[0] == (512 if kind == "view512" else 1920))
                            for _, k, _, v in trace]) & ok
            if sel.any():
                r[f"p99_ms_{kind}"] = float(np.percentile(lat[sel] * 1e3, 99))
        results[name] = r
        log(f"{name}: p50 {r['p50_ms']:.3f} ms p99 {r['p99_ms']:.3f} ms (points {r.get('p99_ms_points', 0):.3f}, "
            f"512^2 {r.get('p99_ms_view512', 0):.3f}, 1080p {r.get('p99_ms_view1080', 0):.3f}); "
            f"{r['mrays_s_at_2k']:.1f} Mrays/s at 2k req/s in {r['rounds']} rounds; saturation "
            f"{r['saturation_mrays_s']:.1f} Mrays/s (median of {r['saturation_walls_s']} s, {r['saturation_rounds']} rounds); "
            f"arena peak {r['arena_peak_mib']:.0f} MiB")
    return results, work, offered, trace


def run_config3(args, ctx, M, L, scenes):
    """Served mixed stream over 8 resident config1-type scenes (seeds 1-8):
    Poisson arrivals at 2000 requests/s (70% point batches U{1k..64k}, 25%
    512x512 views, 5% 1920x1080 views, each view with a shadow ray per hit),
    admitted by a 4 GiB memory budget. The same trace is replayed by a native
    client against the batched server and against the per-request policy
    (one plain launch per request), first at its arrival times (latency), then
    all at once (saturation throughput). Both use the server's default shape
    (4 rounds in flight, rounds of 2^20 rays + queries or a quarter of the
    admitted queue's work if that is more, a round of one request on its
    plain kernel). Point results are copied to pinned host
    memory; view G-buffers stay in HBM for their GPU consumer and are
    released on completion."""
    n_scenes = 8
    resident = [build_scene(ctx, M, scenes, seed) for seed in range(1, n_scenes + 1)]
    ctx.sync()
    gib = sum(float(s.info["device_bytes"]) for s in resident) / 2**30
    log(f"config3: {n_scenes} resident scenes, {sum(int(s.info['n_bricks']) for s in resident)} bricks, {gib:.1f} GiB")
    results, work, offered, trace = served_vs_plain(ctx, M, L, scenes, resident, args.requests)
    ratio = results["served"]["saturation_mrays_s"] / results["plain"]["saturation_mrays_s"]
    line = {
        "metric": "served Mrays/s (rays + queries) at saturation; p50/p99 latency at 2000 req/s, served vs plain",
        "value": round(results["served"]["saturation_mrays_s"], 3), "unit": "Mrays/s", "n_gpus": 1, "steps": 3,
        "warmup": 1, "ms_per_step": round(results["served"]["saturation_s"] * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded)",
        "config": {"workload": "config3: 8 resident config1-type scenes (seeds 1-8); Poisson 2000 req/s; 70% point "
                               "batches U{1k..64k}, 25% 512x512 views, 5% 1920x1080 views (+ shadow ray per hit); "
                               "4 GiB admission budget",
                   "requests": len(trace), "work_items": int(work), "offered_mrays_s": round(offered, 2),
                   "served_vs_plain_saturation": round(ratio, 3),
                   "served": {k: round(v, 4) if isinstance(v, float) else v for k, v in results["served"].items()},
                   "plain": {k: round(v, 4) if isinstance(v, float) else v for k, v in results["plain"].items()}},
    }
    print(json.dumps(line), flush=True)
    for s in resident:
        s.destroy()


def run_config4(args, ctx, M, L, scenes):
    """Terrain stress case: 1 km x 1 km fBm heightfield (6 octaves) at
    vs 0.05 unioned with 20k smooth-min rocks; per frame 32 random sphere
    add/subtract edits (radius U[0.5,3] m, smooth union / smooth subtract,
    blend 0.3 r) rewrite the affected bricks in one launch, then a 3840x2160
    view plus one shadow ray per hit is marched. A step is one frame: edits +
    march. Edits land on the terrain in front of the camera."""
    prims, lo, hi = scenes.terrain_scene(seed=4)
    t0 = time.time()
    sc = M.SparseScene.build(ctx, prims, 0.05, lo, hi, smin_k=0.1, mat_seed=scenes.material_seed(4))
    ctx.sync()
    log(f"config4 scene: {int(sc.info['n_bricks'])} bricks, {int(sc.info['n_blocks'])} L1 blocks, cells "
        f"{[int(c) for c in sc.info['cells']]}, {sc.info['device_bytes'] / 2**30:.2f} GiB, built in {time.time() - t0:.1f} s")
    W, H = 3840, 2160
    t = prims[0]
    cam, su, ext = scenes.terrain_view(t, W, H, cull=not args.no_cull)
    rng = scenes.Pcg32(44)

    def frame_edits():
        return scenes.terrain_edits(t, rng)

    gbd = M.GBuffer(ctx, W, H, debug=True, shadow=True)
    sc.march(cam, su, ext, gbd)
    out = gbd.read()
    hit = out["position"][:, 3] != L.MAX_FLOAT
    nbytes = algorithmic_bytes(out["debug"], hit)
    n_shadow = int(hit.sum())
    dbg = out["debug"]
    log(f"hits {n_shadow} / {W * H}; steps/ray mean {dbg['steps'].mean():.2f} p99 {np.percentile(dbg['steps'], 99):.0f}; "
        f"shadow steps mean {dbg['shadow_steps'][hit].mean() if hit.any() else 0:.1f}; algorithmic bytes {nbytes / 1e9:.3f} GB")
    del gbd, out
    gb = M.GBuffer(ctx, W, H, shadow=True)
    edits = [frame_edits() for _ in range(args.warmup + args.steps)]
    mod = sc.edit(edits[0])
    log(f"first frame's edits modified {len(mod)} bricks")
    for k in range(1, args.warmup):
        sc.edit(edits[k], report=False)
        sc.march(cam, su, ext, gb)
    ctx.sync()
    # frame = edits + march
    ctx.timer_begin()
    for k in range(args.steps):
        sc.edit(edits[args.warmup + k], report=False)
        sc.march(cam, su, ext, gb)
    ms = ctx.timer_end() / args.steps
    ctx.sync()
    ctx.timer_begin()
    for k in range(args.steps):
        sc.march(cam, su, ext, gb)
    ms_march = ctx.timer_end() / args.steps
    ctx.sync()
    ctx.timer_begin()
    for k in range(args.steps):
        sc.edit(edits[k], report=False)
    ms_edit = ctx.timer_end() / args.steps
    ctx.sync()
    rays = W * H + n_shadow
    log(f"config4: frame {ms:.3f} ms (march {ms_march:.3f} ms, 32 edits {ms_edit:.3f} ms)")
    line = {
        "metric": "Mrays/s of the 4K terrain frame with 32 CSG edits per frame",
        "value": round(rays / (ms / 1e3) / 1e6, 3), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded)",
        "config": {"workload": "config4: 1 km^2 fBm terrain (6 octaves, vs 0.05) + 20k smooth-min rocks; 3840x2160 "
                               "primary + 1 shadow ray per hit; 32 sphere add/subtract edits r~U[0.5,3] per frame",
                   "bricks": int(sc.info["n_bricks"]), "scene_gib": round(float(sc.info["device_bytes"]) / 2**30, 2),
                   "primary_rays": W * H, "shadow_rays": n_shadow, "tile_cull": not args.no_cull,
                   "march_ms": round(ms_march, 4),
                   "edits_ms": round(ms_edit, 4),
                   "march_algorithmic_gbs": round(nbytes / (ms_march / 1e3) / 1e9, 2)},
    }
    print(json.dumps(line), flush=True)
    sc.destroy()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="config1", choices=["config1", "config0", "config2", "config3", "config4"])
    ap.add_argument("--requests", type=int, default=4000, help="config3: requests in the arrival trace")
    ap.add_argument("--variant", type=int, default=0,
                    help="sparse kernel: 0 per-pixel, 1 persistent refill, 2 primary + compacted shading pass")
    ap.add_argument("--no-cull", action="store_true",
                    help="config1/config4: march without per-tile ray intervals")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--served-requests", type=int, default=4000,
                    help="config1: requests of the served-vs-plain replay reported beside the frame (the config3 "
                         "trace's 4000 by default; 0 = skip)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also launch the plain copy kernel on this many MiB (profiler counter calibration)")
    ap.add_argument("--gather-mib", type=int, default=0,
                    help="also launch the scattered 2-byte read kernel over this many MiB, one read per 128-B line "
                         "(profiler counter calibration for gathers)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the path shards rays across GPUs with no exchange, "
              "and runs on more than one GPU are not part of this build", file=sys.stderr)
        sys.exit(2)

    from rkifield_amd import layouts as L
    from rkifield_amd import march as M
    from rkifield_amd import scenes

    ctx = M.Context(0)
    if args.workload in ("config2", "config3", "config4"):
        {"config2": run_config2, "config3": run_config3, "config4": run_config4}[args.workload](args, ctx, M, L, scenes)
        ctx.close()
        return
    if args.workload == "config1":
        width, height = 1920, 1080
        sc, cam, su, ext = build_config1(ctx, M, L, scenes, width, height)
        ext["march_flags"] = 0 if args.no_cull else L.MARCH_TILE_CULL
        launch = lambda gb: sc.march(cam, su, ext, gb, variant=args.variant)
        workload = ("config1: 4096 sphere/box/capsule prims, size U[0.1,4] m, smin k=0.1 groups of 8, "
                    "256 m cube, vs 0.02 m two-level sparse index; 1920x1080 primary + 1 shadow ray per hit")
        kernel = ["sparse_march_plain", "sparse_march_tuned", "sparse_march_split"][args.variant]
    else:
        sc = scenes.spheres_scene(seed=0)
        width, height = sc.width, sc.height
        gs = M.GpuScene(ctx)
        gs.upload_brick_pool(sc.pool)
        gs.upload_brick_maps(sc.maps)
        gs.upload_objects(sc.objects)
        gs.update_camera(sc.cam)
        gs.update_scene(sc.su, sc.ext)
        tc = M.TileObjectCullPass(ctx, width, height)
        tc.dispatch()
        cf = M.CoarseField.from_scene_aabbs(ctx, sc.aabbs())
        cf.populate_and_upload()
        rp = M.RayMarchPass(ctx)
        launch = lambda gb: rp.dispatch(gs, gb, tc, cf)
        workload = "config0: 64 voxelised spheres (seed 0), vs 0.05 m, reference layouts, 640x360 primary rays"
        kernel = "ray_march_ref_kernel"

    # counting launch (debug plane on): evaluation counts -> algorithmic bytes
    gbd = M.GBuffer(ctx, width, height, debug=True, shadow=args.workload == "config1")
    launch(gbd)
    out = gbd.read()
    hit = out["position"][:, 3] != L.MAX_FLOAT
    nbytes = algorithmic_bytes(out["debug"], hit)
    n_primary = width * height
    n_shadow = int(hit.sum()) if args.workload == "config1" else 0
    dbg = out["debug"]
    log(f"hits {int(hit.sum())} / {n_primary}; evals {int(dbg['evals'].sum())}, "
        f"shadow evals {int(dbg['shadow_evals'].sum())}; algorithmic bytes {nbytes / 1e9:.3f} GB")
    log(f"steps/ray mean {dbg['steps'].mean():.2f} p99 {np.percentile(dbg['steps'], 99):.0f} max {dbg['steps'].max()}; "
        f"evals/ray mean {dbg['evals'].mean():.2f}; hit rays: steps mean {dbg['steps'][hit].mean() if hit.any() else 0:.1f}; "
        f"shadow steps mean {dbg['shadow_steps'][hit].mean() if hit.any() else 0:.1f}")
    del gbd

    gb = M.GBuffer(ctx, width, height, shadow=args.workload == "config1")
    for _ in range(args.warmup):
        launch(gb)
    ctx.sync()
    ctx.timer_begin()
    for _ in range(args.steps):
        launch(gb)
    total_ms = ctx.timer_end()
    ctx.sync()
    ms = total_ms / args.steps
    rays = n_primary + n_shadow
    mrays = rays / (ms / 1e3) / 1e6
    achieved = nbytes / (ms / 1e3) / 1e9
    log(f"{kernel}: {ms:.3f} ms/launch, {mrays:.1f} Mrays/s, {achieved:.1f} GB/s algorithmic")

    if args.copy_mib:
        nb = args.copy_mib << 20
        a = ctx.array((nb,), np.uint8)
        b = ctx.array((nb,), np.uint8)
        for _ in range(3):
            ctx.copy(b, a, nb)
        ctx.sync()
        log(f"copy kernel: 3 x {args.copy_mib} MiB")
    if args.gather_mib:
        nb = args.gather_mib << 20
        src = ctx.array((nb,), np.uint8)
        n = nb // 128
        dst = ctx.array((n,), np.uint16)
        for _ in range(3):
            ctx.gather16(dst, src, nb, n)
        ctx.sync()
        log(f"gather16 kernel: 3 x {n} lines of {args.gather_mib} MiB")

    traffic = None
    tpath =os.path.join(ROOT, "profiles", f"{kernel}_{args.workload}_traffic.json")
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get("bytes_per_launch")
        except (OSError, ValueError):
            traffic = None

    cpu = acc = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        if args.workload == "config1":
            cpu, acc = cpu_baseline_sparse(sc, cam, su, ext, width, height, args.cpu_seconds, gpu=out)
        else:
            idx, counts = tc.read()
            cells, u = cf.read()
            cpu, twin = cpu_baseline_ref(sc, sc.ext, idx, counts, cells, u, args.cpu_seconds)
            acc = accuracy_ref(sc, sc.ext, idx, counts, cells, u, out, twin)
        log(f"cpu baseline {cpu['value']:.3f} Mrays/s on {cpu['cores']} threads; accuracy {acc}")

    served = None
    if args.workload == "config1" and args.served_requests > 0:
        others = [build_scene(ctx, M, scenes, seed) for seed in range(2, 9)]
        ctx.sync()
        res_s, work_s, offered, _ = served_vs_plain(ctx, M, L, scenes, [sc] + others, args.served_requests)
        keys = ("p50_ms", "p99_ms", "p99_ms_points", "p99_ms_view512", "p99_ms_view1080", "mrays_s_at_2k",
                "saturation_mrays_s", "failed")
        served = {"workload": f"{args.served_requests} requests of the served trace (Poisson 2000 req/s; 70% point "
                              f"batches U{{1k..64k}}, 25% 512x512, 5% 1920x1080 views + shadow rays) over 8 "
                              f"resident scenes of this kind (seeds 1-8), 4 GiB admission budget",
                  "offered_mrays_s": round(offered, 2),
                  "served": {k: round(res_s["served"][k], 4) for k in keys if k in res_s["served"]},
                  "plain": {k: round(res_s["plain"][k], 4) for k in keys if k in res_s["plain"]},
                  "served_vs_plain_saturation": round(res_s["served"]["saturation_mrays_s"]
                                                      / res_s["plain"]["saturation_mrays_s"], 3)}
        for o in others:
            o.destroy()

    line = {
        "metric": "Mrays/s + % HBM roofline at 1080p; p50/p99 latency served vs plain",
        "value": round(mrays, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded splitmix64/PCG32 scene and camera)",
        "config": {"workload": workload, "kernel": kernel, "width": width, "height": height,
                   "primary_rays": n_primary, "shadow_rays": n_shadow,
                   "tile_cull": args.workload == "config1" and not args.no_cull},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic},
        "cpu_baseline": cpu,
        "accuracy": acc,
        "served": served,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()

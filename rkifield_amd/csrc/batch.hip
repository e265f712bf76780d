// batch.hip — one launch over many uneven batches (views, shadow-ray batches,
// point-query batches, sphere-cast batches), possibly over different resident
// sparse scenes.
//
// Work units: a view is cut into tile_w x tile_h pixel packets (64 lanes), a
// shadow or point batch into packets of 64. Unit u belongs to the batch whose
// [unit_begin, unit_begin + n_units) holds it (binary search over <= a few
// hundred wave-uniform begins, through the scalar cache). Each unit runs the
// same per-ray code as the single-batch entry points (view_pixel,
// shadow_step, point_query, sphere_cast), so every result equals the batch
// run alone.
//
// Schedules: GRID launches one wavefront per unit and lets the hardware
// dispatcher balance; PERSISTENT launches waves_per_cu x CUs wavefronts that
// pull units from a counter (one atomic per 64 rays); GRID_LOWLAT is GRID
// with the single-view kernels' per-ray code (brick-neighbourhood cache),
// which shortens the slowest rays at the price of more memory instructions.
// The autotuner times them and the packet shapes on the actual batch set.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "host_internal.h"
#include "sparse_device.h"

namespace rkf {

RKF_DEV uint32_t find_batch(const uint32_t *begins, uint32_t n_items, uint32_t unit) {
    uint32_t lo = 0, hi = n_items;
    while (hi - lo > 1u) {
        uint32_t mid = (lo + hi) >> 1;
        if (begins[mid] <= unit) lo = mid;
        else hi = mid;
    }
    return lo;
}

template <bool kLowLat>
RKF_DEV void run_unit(const BatchDev &B, uint32_t u, uint32_t lane) {
    if (B.kind == RKFM_BATCH_VIEW) {
        uint32_t tx = u % B.tiles_x, ty = u / B.tiles_x;
        uint32_t px = tx * B.tile_w + lane % B.tile_w;
        uint32_t py = ty * B.tile_h + lane / B.tile_w;
        view_pixel<kLowLat>(B.P, px, py, px < B.P.g.width && py < B.P.g.height);
    } else if (B.kind == RKFM_BATCH_SHADOW) {
        uint32_t i = u * 64u + lane;
        if (i >= B.n) return;
        f3 o = mk(B.in[3 * (uint64_t)i], B.in[3 * (uint64_t)i + 1], B.in[3 * (uint64_t)i + 2]);
        ShadowRay sr;
        shadow_init(sr, o, mk(B.P.ext.light_dir[0], B.P.ext.light_dir[1], B.P.ext.light_dir[2]), B.P.S.vs);
        float maxd = shadow_maxd(B.P), k = shadow_k(B.P);
        while (!shadow_step<kLowLat>(B.P.S, sr, maxd, k)) {
        }
        B.vis[i] = sr.prev_dist;
    } else if (B.kind == RKFM_BATCH_POINTS) {
        uint32_t i = u * 64u + lane;
        if (i >= B.n) return;
        f3 p = mk(B.in[3 * (uint64_t)i], B.in[3 * (uint64_t)i + 1], B.in[3 * (uint64_t)i + 2]);
        point_query(B.P.S, p, B.res[i]);
    } else {
        uint32_t i = u * 64u + lane;
        if (i >= B.n) return;
        sphere_cast(B.P.S, B.casts[i], B.cres[i]);
    }
}

// The batch's scene pointers as global pointers (they come from the
// descriptor array, not from kernel arguments).
RKF_DEV void global_scene(SparseDev &S) {
    S.l2 = global_ptr(S.l2);
    S.l1 = global_ptr(S.l1);
    S.dist = global_ptr(S.dist);
    S.mat = global_ptr(S.mat);
    S.aux = global_ptr(S.aux);
    S.free_idx = global_ptr(S.free_idx);
    S.free_bits = global_ptr(S.free_bits);
    S.oct = global_ptr(S.oct);
    S.roww = global_ptr(S.roww);
    S.cellw = global_ptr(S.cellw);
    S.block_cell = global_ptr(S.block_cell);
}

// Workgroups go to the 8 XCDs round-robin (workgroup b to XCD b % 8): each
// run of 8 x 4 workgroups hands XCD x the 4 consecutive units 4 x .. 4 x + 3
// of the run (a view's neighbouring tiles of a row share bricks in that
// XCD's L2; config 2: 1.74 -> 1.68 ms, runs of 2 / 8 / 16: 1.70 / 1.68 /
// 1.71). A last partial run keeps the plain mapping.
constexpr uint32_t kXcds = 8u, kXcdRun = 4u;
template <bool kLowLat>
__global__ __launch_bounds__(64) void batch_grid_kernel(const BatchDev *__restrict__ items, const uint32_t *__restrict__ begins,
                                                        uint32_t n_items, uint32_t total) {
    uint32_t unit = blockIdx.y * gridDim.x + blockIdx.x;
    const uint32_t run = unit / (kXcds * kXcdRun), in_run = unit % (kXcds * kXcdRun);
    if ((run + 1u) * kXcds * kXcdRun <= gridDim.x * gridDim.y)
        unit = run * kXcds * kXcdRun + (in_run % kXcds) * kXcdRun + in_run / kXcds;
    if (unit >= total) return;
    uint32_t b = find_batch(begins, n_items, unit);
    BatchDev B = items[b];
    global_scene(B.P.S);
    run_unit<kLowLat>(B, unit - B.unit_begin, threadIdx.x);
}

__global__ __launch_bounds__(64) void batch_persistent_kernel(const BatchDev *__restrict__ items, const uint32_t *__restrict__ begins,
                                                              uint32_t n_items, uint32_t total, uint32_t *counter) {
    const uint32_t lane = threadIdx.x;
    while (true) {
        uint32_t unit = 0;
        if (lane == 0) unit = atomicAdd(counter, 1u);
        unit = __builtin_amdgcn_readfirstlane(unit);  // lane 0's (all lanes are active here)
        if (unit >= total) break;
        uint32_t b = find_batch(begins, n_items, unit);
        BatchDev B = items[b];
        global_scene(B.P.S);
        run_unit<false>(B, unit - B.unit_begin, lane);
    }
}

__global__ void batch_reset_kernel(uint32_t *counter) { *counter = 0u; }

}  // namespace rkf

struct rkfm_batch_graph {
    rkfm_ctx *ctx = nullptr;
    DevBuf items, begins, counter, tiles;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // Recorded on the stream of every replay: destroy waits for it, so the
    // buffers above outlive the last replay whatever stream it ran on.
    hipEvent_t done = nullptr;
};

namespace {

bool owned(const rkfm_ctx *c, const rkfm_sparse_scene *s) {
    return std::find(c->scenes.begin(), c->scenes.end(), s) != c->scenes.end();
}

}  // namespace

bool rkf_batch_valid_shape(const rkfm_batch_tuning &t) {
    if (t.schedule > RKFM_SCHEDULE_GRID_LOWLAT) return false;
    if (t.tile_w == 0 || t.tile_h == 0 || t.tile_w * t.tile_h != 64u) return false;
    if (t.schedule == RKFM_SCHEDULE_PERSISTENT && (t.waves_per_cu == 0 || t.waves_per_cu > 64)) return false;
    return true;
}

// Host descriptors for a batch set; returns the unit total. c == nullptr
// skips the check that every scene is resident in the context (the caller
// has checked it).
int rkf_batch_prepare(rkfm_ctx *c, const rkfm_batch_item *items, uint32_t n, const rkfm_batch_tuning &t,
            std::vector<rkf::BatchDev> &dev, std::vector<uint32_t> &begins, uint32_t &total,
            std::vector<uint32_t> *cull_views) {
    if (cull_views) cull_views->clear();
    if (!items || n == 0) return rkf_set_error(RKFM_ERR_INVALID_ARG, "batch: no items");
    if (!rkf_batch_valid_shape(t)) return rkf_set_error(RKFM_ERR_INVALID_ARG, "batch: invalid launch shape");
    dev.assign(n, rkf::BatchDev{});
    begins.assign(n, 0);
    uint64_t acc = 0;
    for (uint32_t i = 0; i < n; i++) {
        const rkfm_batch_item &it = items[i];
        rkf::BatchDev &d = dev[i];
        memset(&d, 0, sizeof(d));
        if (!it.scene || (c && !owned(c, it.scene)))
            return rkf_set_error(RKFM_ERR_INVALID_ARG, "batch item %u: scene not resident in this context", i);
        if (it.ext.hit_eps_mode > RKFM_EPS_RELATIVE)
            return rkf_set_error(RKFM_ERR_INVALID_ARG, "batch item %u: hit_eps_mode", i);
        const bool cull = it.ext.march_flags == RKFM_MARCH_TILE_CULL && it.kind == RKFM_BATCH_VIEW && cull_views;
        if (it.ext.march_flags && !cull)
            return rkf_set_error(RKFM_ERR_INVALID_ARG, "batch item %u: march_flags 0x%x not supported here", i,
                                 it.ext.march_flags);
        d.P.S = it.scene->dev();
        d.P.cam = it.camera;
        d.P.su = it.su;
        d.P.ext = it.ext;
        d.P.object_id = 0;
        d.kind = it.kind;
        d.n = it.n;
        d.tile_w = t.tile_w;
        d.tile_h = t.tile_h;
        uint64_t units = 0;
        if (it.kind == RKFM_BATCH_VIEW) {
            int rc = rkf_fill_gbuf(&it.out, d.P.g);
            if (rc) return rkf_set_error(rc, "batch item %u: %s", i, rkfm_last_error_string());
            d.tiles_x = (it.out.width + t.tile_w - 1) / t.tile_w;
            units = (uint64_t)d.tiles_x * ((it.out.height + t.tile_h - 1) / t.tile_h);
            if (cull) {
                rc = rkf_tile_camera_check(&it.camera);
                if (rc) return rkf_set_error(rc, "batch item %u: %s", i, rkfm_last_error_string());
                cull_views->push_back(i);
            }
        } else if (it.kind == RKFM_BATCH_SHADOW) {
            if (it.n && (!it.origins || !it.visibility))
                return rkf_set_error(RKFM_ERR_INVALID_ARG, "batch item %u: null shadow buffers", i);
            d.in = it.origins;
            d.vis = it.visibility;
            units = (it.n + 63u) / 64u;
        } else if (it.kind == RKFM_BATCH_POINTS) {
            if (it.n && (!it.xyz || !it.results))
                return rkf_set_error(RKFM_ERR_INVALID_ARG, "batch item %u: null point buffers", i);
            d.in = it.xyz;
            d.res = it.results;
            units = (it.n + 63u) / 64u;
        } else if (it.kind == RKFM_BATCH_SPHERES) {
            if (it.n && (!it.casts || !it.cast_results))
                return rkf_set_error(RKFM_ERR_INVALID_ARG, "batch item %u: null sphere-cast buffers", i);
            d.casts = it.casts;
            d.cres = it.cast_results;
            units = (it.n + 63u) / 64u;
        } else {
            return rkf_set_error(RKFM_ERR_INVALID_ARG, "batch item %u: kind %u", i, it.kind);
        }
        d.unit_begin = (uint32_t)acc;
        d.n_units = (uint32_t)units;
        begins[i] = (uint32_t)acc;
        acc += units;
        if (acc >= 0xFFFFFFF0ull) return rkf_set_error(RKFM_ERR_INVALID_ARG, "batch: too many work units");
    }
    total = (uint32_t)acc;
    return RKFM_OK;
}

int rkf_device_waves(uint32_t per_cu, uint32_t &waves) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return rkf_hip_check(e, "hipGetDevice");
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, dev);
    if (e != hipSuccess) return rkf_hip_check(e, "hipGetDeviceProperties");
    waves = per_cu * (uint32_t)prop.multiProcessorCount;
    return RKFM_OK;
}

// Enqueues reset (persistent only) + kernel on `st`; `waves` is the
// persistent grid (rkf_device_waves, computed outside any capture).
int rkf_batch_enqueue(const rkfm_batch_tuning &t, const rkf::BatchDev *items, const uint32_t *begins, uint32_t n,
            uint32_t total, uint32_t *counter, uint32_t waves, hipStream_t st) {
    if (total == 0) return RKFM_OK;
    if (t.schedule == RKFM_SCHEDULE_GRID || t.schedule == RKFM_SCHEDULE_GRID_LOWLAT) {
        uint32_t gx = total < 65536u ? total : 65536u;
        uint32_t gy = (total + gx - 1) / gx;
        if (t.schedule == RKFM_SCHEDULE_GRID)
            hipLaunchKernelGGL(rkf::batch_grid_kernel<false>, dim3(gx, gy), dim3(64), 0, st, items, begins, n, total);
        else
            hipLaunchKernelGGL(rkf::batch_grid_kernel<true>, dim3(gx, gy), dim3(64), 0, st, items, begins, n, total);
    } else {
        if (waves > total) waves = total;
        if (waves == 0) waves = 1;
        hipLaunchKernelGGL(rkf::batch_reset_kernel, dim3(1), dim3(1), 0, st, counter);
        hipLaunchKernelGGL(rkf::batch_persistent_kernel, dim3(waves), dim3(64), 0, st, items, begins, n, total,
                           counter);
    }
    return rkf_hip_check(hipGetLastError(), "batch launch");
}

rkfm_batch_tuning rkf_batch_default_tuning(const rkfm_ctx *c) {
    if (c->batch_tuned) return c->batch_tuning;
    rkfm_batch_tuning t;
    t.schedule = RKFM_SCHEDULE_GRID;
    t.tile_w = 8;
    t.tile_h = 8;
    t.waves_per_cu = 16;
    return t;
}

namespace {

// Device layout of the tile intervals of a batch's culled views: each view's
// (t_lo, t_hi) per 8x8 tile, then one large-rectangle list and its counter,
// which the views share (they are binned one after another on one stream).
struct CullLayout {
    std::vector<uint64_t> off;
    uint64_t big_off = 0, count_off = 0, bytes = 0;
};

CullLayout cull_layout(const rkfm_batch_item *items, const std::vector<uint32_t> &views) {
    CullLayout L;
    uint64_t acc = 0, nb = 0;
    for (uint32_t i : views) {
        const rkfm_batch_item &it = items[i];
        L.off.push_back(acc);
        acc += (uint64_t)((it.out.width + 7u) / 8u) * ((it.out.height + 7u) / 8u) * sizeof(uint2);
        nb = std::max<uint64_t>(nb, it.scene->info.n_blocks);
    }
    L.big_off = (acc + 255u) & ~255ull;
    L.count_off = L.big_off + ((nb * sizeof(rkf::TileBig) + 64u + 255u) & ~255ull);
    L.bytes = L.count_off + 256u;
    return L;
}

// Points the culled views' descriptors at their tile arrays under `base`.
void cull_patch(std::vector<rkf::BatchDev> &dev, const rkfm_batch_item *items, const std::vector<uint32_t> &views,
                const CullLayout &L, unsigned char *base) {
    for (size_t k = 0; k < views.size(); k++) {
        rkf::BatchDev &d = dev[views[k]];
        d.P.tiles = (const uint2 *)(base + L.off[k]);
        d.P.tiles_x = (items[views[k]].out.width + 7u) / 8u;
    }
}

// The tile-interval kernels of every culled view (the plain path's own
// launches, so a culled view marches exactly as it does alone).
int cull_enqueue(const rkfm_batch_item *items, const std::vector<uint32_t> &views, const CullLayout &L,
                 unsigned char *base, hipStream_t st) {
    for (size_t k = 0; k < views.size(); k++) {
        const rkfm_batch_item &it = items[views[k]];
        const rkf::SparseDev D = it.scene->dev();
        hipError_t e = rkf_launch_sparse_tiles(&D, &it.camera, it.su.max_distance, it.out.width, it.out.height,
                                               (uint2 *)(base + L.off[k]), nullptr, (rkf::TileBig *)(base + L.big_off),
                                               (uint32_t *)(base + L.count_off), true, st);
        if (e != hipSuccess) return rkf_hip_check(e, "batch tile intervals");
    }
    return RKFM_OK;
}

}  // namespace

extern "C" {

uint64_t rkfm_batch_work_items(const rkfm_batch_item *items, uint32_t n) {
    uint64_t s = 0;
    for (uint32_t i = 0; items && i < n; i++)
        s += items[i].kind == RKFM_BATCH_VIEW ? (uint64_t)items[i].out.width * items[i].out.height : items[i].n;
    return s;
}

int rkfm_batch_run(rkfm_ctx *c, const rkfm_batch_item *items, uint32_t n, const rkfm_batch_tuning *tuning,
                   void *stream) {
    if (!c) return rkf_set_error(RKFM_ERR_INVALID_ARG, "batch_run: null ctx");
    rkfm_batch_tuning t = tuning ? *tuning : rkf_batch_default_tuning(c);
    std::vector<rkf::BatchDev> dev;
    std::vector<uint32_t> begins, cull;
    uint32_t total = 0;
    int rc = rkf_batch_prepare(c, items, n, t, dev, begins, total, &cull);
    if (rc) return rc;
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    CullLayout cl;
    if (!rc && !cull.empty()) {
        cl = cull_layout(items, cull);
        rc = c->batch_tiles.reserve(cl.bytes, false);
        if (rc) return rc;
        cull_patch(dev, items, cull, cl, (unsigned char *)c->batch_tiles.ptr);
    }
    // The descriptors are rewritten per call: order against earlier launches
    // that read them (they ran on this stream or were synchronised).
    uint32_t waves = 0;
    if (t.schedule == RKFM_SCHEDULE_PERSISTENT) {
        rc = rkf_device_waves(t.waves_per_cu, waves);
        if (rc) return rc;
    }
    const uint64_t ibytes = dev.size() * sizeof(rkf::BatchDev);
    rc = c->batch_items.reserve(ibytes, false);
    if (!rc) rc = c->batch_begins.reserve(begins.size() * 4, false);
    if (!rc) rc = c->batch_counter.reserve(256, false);
    if (rc) return rc;
    c->batch_host_items.resize(ibytes);
    memcpy(c->batch_host_items.data(), dev.data(), ibytes);
    c->batch_host_begins.swap(begins);
    rc = rkf_hip_check(hipMemcpyAsync(c->batch_items.ptr, c->batch_host_items.data(), ibytes, hipMemcpyHostToDevice,
                                      st),
                       "batch descriptors");
    if (!rc)
        rc = rkf_hip_check(hipMemcpyAsync(c->batch_begins.ptr, c->batch_host_begins.data(),
                                          c->batch_host_begins.size() * 4, hipMemcpyHostToDevice, st),
                           "batch begins");
    if (!rc && !cull.empty()) rc = cull_enqueue(items, cull, cl, (unsigned char *)c->batch_tiles.ptr, st);
    if (rc) return rc;
    return rkf_batch_enqueue(t, (const rkf::BatchDev *)c->batch_items.ptr, (const uint32_t *)c->batch_begins.ptr, n, total,
                   (uint32_t *)c->batch_counter.ptr, waves, st);
}

int rkfm_batch_autotune(rkfm_ctx *c, const rkfm_batch_item *items, uint32_t n, uint32_t reps,
                        rkfm_batch_tuning *best, float *best_ms) {
    if (!c || !best) return rkf_set_error(RKFM_ERR_INVALID_ARG, "batch_autotune: null argument");
    if (reps == 0) reps = 3;
    static const uint32_t shapes[][2] = {{8, 8}, {16, 4}, {4, 16}, {32, 2}};
    std::vector<rkfm_batch_tuning> cands;
    for (auto &s : shapes) cands.push_back({RKFM_SCHEDULE_GRID, s[0], s[1], 0});
    for (int k = 0; k < 2; k++) cands.push_back({RKFM_SCHEDULE_GRID_LOWLAT, shapes[k][0], shapes[k][1], 0});
    for (uint32_t w : {8u, 16u, 24u})
        for (int k = 0; k < 2; k++) cands.push_back({RKFM_SCHEDULE_PERSISTENT, shapes[k][0], shapes[k][1], w});
    float best_t = 3.4e38f;
    rkfm_batch_tuning bt = cands[0];
    for (const rkfm_batch_tuning &t : cands) {
        int rc = rkfm_batch_run(c, items, n, &t, nullptr);  // warm-up
        if (rc) return rc;
        rc = rkf_hip_check(hipEventRecord(c->ev0, c->stream), "event");
        for (uint32_t r = 0; !rc && r < reps; r++) rc = rkfm_batch_run(c, items, n, &t, nullptr);
        if (!rc) rc = rkf_hip_check(hipEventRecord(c->ev1, c->stream), "event");
        if (!rc) rc = rkf_hip_check(hipEventSynchronize(c->ev1), "event sync");
        float ms = 0.0f;
        if (!rc) rc = rkf_hip_check(hipEventElapsedTime(&ms, c->ev0, c->ev1), "elapsed");
        if (rc) return rc;
        ms /= (float)reps;
        if (ms < best_t) {
            best_t = ms;
            bt = t;
        }
    }
    *best = bt;
    if (best_ms) *best_ms = best_t;
    c->batch_tuning = bt;
    c->batch_tuned = true;
    return RKFM_OK;
}

int rkfm_batch_graph_create(rkfm_ctx *c, const rkfm_batch_item *items, uint32_t n, const rkfm_batch_tuning *tuning,
                            rkfm_batch_graph **out) {
    if (!c || !out) return rkf_set_error(RKFM_ERR_INVALID_ARG, "batch_graph_create: null argument");
    *out = nullptr;
    rkfm_batch_tuning t = tuning ? *tuning : rkf_batch_default_tuning(c);
    std::vector<rkf::BatchDev> dev;
    std::vector<uint32_t> begins, cull;
    uint32_t total = 0;
    int rc = rkf_batch_prepare(c, items, n, t, dev, begins, total, &cull);
    if (rc) return rc;
    uint32_t waves = 0;
    if (t.schedule == RKFM_SCHEDULE_PERSISTENT) {
        rc = rkf_device_waves(t.waves_per_cu, waves);
        if (rc) return rc;
    }
    rkfm_batch_graph *g = new rkfm_batch_graph();
    g->ctx = c;
    rc = rkf_hip_check(hipEventCreateWithFlags(&g->done, hipEventDisableTiming), "graph event");
    CullLayout cl;
    if (!rc && !cull.empty()) {
        cl = cull_layout(items, cull);
        rc = g->tiles.reserve(cl.bytes, false);
        if (!rc) cull_patch(dev, items, cull, cl, (unsigned char *)g->tiles.ptr);
    }
    if (!rc) rc = g->items.reserve(dev.size() * sizeof(rkf::BatchDev), false);
    if (!rc) rc = g->begins.reserve(begins.size() * 4, false);
    if (!rc) rc = g->counter.reserve(256, false);
    if (!rc)
        rc = rkf_hip_check(hipMemcpy(g->items.ptr, dev.data(), dev.size() * sizeof(rkf::BatchDev),
                                     hipMemcpyHostToDevice),
                           "graph descriptors");
    if (!rc) rc = rkf_hip_check(hipMemcpy(g->begins.ptr, begins.data(), begins.size() * 4, hipMemcpyHostToDevice),
                                "graph begins");
    hipStream_t cap = nullptr;
    if (!rc) rc = rkf_hip_check(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking), "capture stream");
    if (!rc) rc = rkf_hip_check(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal), "begin capture");
    if (!rc) {
        int rc2 = cull.empty() ? RKFM_OK : cull_enqueue(items, cull, cl, (unsigned char *)g->tiles.ptr, cap);
        if (!rc2)
            rc2 = rkf_batch_enqueue(t, (const rkf::BatchDev *)g->items.ptr, (const uint32_t *)g->begins.ptr, n, total,
                                    (uint32_t *)g->counter.ptr, waves, cap);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(cap, &graph);
        rc = rc2 ? rc2 : rkf_hip_check(e, "end capture");
        g->graph = graph;
    }
    if (!rc) rc = rkf_hip_check(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0), "graph instantiate");
    if (cap) (void)hipStreamDestroy(cap);
    if (rc) {
        rkfm_batch_graph_destroy(g);
        return rc;
    }
    *out = g;
    return RKFM_OK;
}

int rkfm_batch_graph_launch(rkfm_batch_graph *g, void *stream) {
    if (!g || !g->exec) return rkf_set_error(RKFM_ERR_INVALID_ARG, "batch_graph_launch: null graph");
    hipStream_t st = stream ? (hipStream_t)stream : g->ctx->stream;
    int rc = rkf_hip_check(hipGraphLaunch(g->exec, st), "graph launch");
    if (!rc) rc = rkf_hip_check(hipEventRecord(g->done, st), "graph launch event");
    return rc;
}

int rkfm_batch_graph_destroy(rkfm_batch_graph *g) {
    if (!g) return RKFM_OK;
    if (g->ctx) (void)hipStreamSynchronize(g->ctx->stream);
    if (g->done) {
        (void)hipEventSynchronize(g->done);  // the last replay, on whichever stream it ran
        (void)hipEventDestroy(g->done);
    }
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    g->items.release();
    g->begins.release();
    g->counter.release();
    g->tiles.release();
    delete g;
    return RKFM_OK;
}

}  // extern "C"

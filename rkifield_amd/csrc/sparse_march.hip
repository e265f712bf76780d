// sparse_march.hip — single-view kernels over a resident sparse scene and the
// point-query kernel. The per-ray arithmetic lives in sparse_device.h.
#include "sparse_device.h"

namespace rkf {

// Plain variant: one lane per pixel, 8x8 pixels per wavefront, each ray
// marched to completion, then its shadow ray. With a tile order (culled
// views) the tiles with most occupied cells start first, so the long ones
// do not start last; without one, tiles go in row-major order. Workgroups go
// to the 8 XCDs round-robin (workgroup b to XCD b % 8), so each run of
// 8 x 16 workgroups hands XCD x the 16 consecutive slots 16 x .. 16 x + 15
// of the run: 16 neighbouring tiles of a row, or tiles the sort wrote side
// by side (often on the same rows), which share bricks in that XCD's L2. A
// last partial run keeps the plain mapping.
constexpr uint32_t kXcds = 8u, kXcdRun = 16u;
// bins of the tile order's counting sort (counts capped at 255)
constexpr uint32_t kOrderBins = 256u;
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void sparse_march_plain(SparseMarchParams P) {
    uint32_t b = blockIdx.y * gridDim.x + blockIdx.x;
    const uint32_t run = b / (kXcds * kXcdRun), in_run = b % (kXcds * kXcdRun);
    if ((run + 1u) * kXcds * kXcdRun <= gridDim.x * gridDim.y)
        b = run * kXcds * kXcdRun + (in_run % kXcds) * kXcdRun + in_run / kXcds;
    const uint32_t tile = P.order ? P.order[b] : b;
    const uint32_t bx = tile % gridDim.x, by = tile / gridDim.x;
    uint32_t px = bx * 8u + (threadIdx.x & 7u);
    uint32_t py = by * 8u + (threadIdx.x >> 3);
    view_pixel<true>(P, px, py, px < P.g.width && py < P.g.height);
    // the tile's interval was read (and used) above; nothing else reads it
    if (P.reset_iv && threadIdx.x == 0u) {
        P.reset_iv[tile] = make_uint2(0x7F800000u, 0u);
        if (P.reset_cnt) P.reset_cnt[tile] = 0u;
    }
    if (P.reset_big && blockIdx.x == 0u && blockIdx.y == 0u) {
        if (threadIdx.x == 0u) *P.reset_big = 0u;
        if (P.reset_bins)
            for (uint32_t i = threadIdx.x; i < 2u * kOrderBins; i += 64u) P.reset_bins[i] = 0u;
    }
}

// Tuned variant: persistent wavefronts with per-lane ray refill and no
// atomics. Wave w of G owns the 8x8 tiles w, w+G, w+2G, ... (tile-major ray
// ids, so a refill takes the next rays of the same tile) and hands them to
// its lanes from a wave-uniform cursor: a lane whose ray finished takes the
// next ray as soon as a quarter of the wave is idle. A lane whose primary ray
// hits turns into that pixel's shadow ray, so shadow marching fills lanes
// instead of stalling the wave. Both ray kinds share one step: position,
// free-space test, trilinear sample; only the bookkeeping differs. Each ray
// runs exactly the arithmetic of prim_step / shadow_step.
constexpr uint32_t kRefillMin = 16u;
constexpr int kInnerSteps = 4;

__global__ __launch_bounds__(64) void sparse_march_tuned(SparseMarchParams P, uint32_t n_waves) {
    const uint32_t lane = threadIdx.x;
    const uint32_t W = P.g.width, H = P.g.height;
    const uint32_t tiles_x = (W + 7u) / 8u;
    const uint32_t n_tiles = tiles_x * ((H + 7u) / 8u);
    const uint32_t wave = blockIdx.x;
    const bool shadows = P.g.shadow && P.ext.shadow_enable;
    const float maxd = shadow_maxd(P), ksh = shadow_k(P);
    const f3 L = mk(P.ext.light_dir[0], P.ext.light_dir[1], P.ext.light_dir[2]);
    // per-wave stream: ray s of this wave is tile (wave + (s / 64) * n_waves), slot s % 64
    uint32_t cursor = 0;
    bool exhausted = wave >= n_tiles;
    // lane state: 0 idle, 1 primary, 2 shadow
    uint32_t kind = 0;
    PrimRay r;
    uint32_t px = 0, py = 0;
    while (true) {
        uint64_t idle = __ballot(kind == 0);
        if (!exhausted && (uint32_t)__popcll(idle) >= kRefillMin) {
            uint32_t n = (uint32_t)__popcll(idle);
            if (kind == 0) {
                uint32_t sidx = cursor + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
                uint32_t tile = wave + (sidx >> 6) * n_waves;
                if (tile < n_tiles) {
                    uint32_t k = sidx & 63u;
                    px = (tile % tiles_x) * 8u + (k & 7u);
                    py = (tile / tiles_x) * 8u + (k >> 3);
                    if (px < W && py < H) {
                        f3 o, dir;
                        camera_ray(P.cam, W, H, px, py, o, dir);
                        prim_init(r, o, dir, P.S.vs, P.su.max_distance);
                        if (P.tiles) prim_cull(r, P.tiles[(uint64_t)(py >> 3) * P.tiles_x + (px >> 3)], P.S.vs);
                        kind = 1;
                    }
                }
            }
            cursor += n;
            if (wave + (cursor >> 6) * n_waves >= n_tiles) exhausted = true;
        }
        if (__ballot(kind != 0) == 0ull) {
            if (exhausted) break;
            continue;
        }
        for (int it = 0; it < kInnerSteps && kind != 0; it++) {
            if (kind == 1) {
                SpRes res;
                int st = prim_step<true>(P.S, r, P.su, P.ext.hit_eps_mode, res);
                if (st != kRayContinue) {
                    f3 so;
                    bool sh = prim_finish(P, px, py, r, st, res, so);
                    if (sh && shadows) {
                        shadow_init(r, so, L, P.S.vs);
                        kind = 2;
                    } else {
                        kind = 0;
                    }
                }
            } else {
                if (shadow_step<true>(P.S, r, maxd, ksh)) {
                    shadow_store(P, (uint64_t)py * W + px, r);
                    kind = 0;
                }
            }
        }
    }
}

// Split variant. Primary pass: one lane per pixel, 8x8 pixels per wavefront;
// misses are written at once, hits are appended (one atomic per wavefront,
// ballot-ranked) to a compact list. Shading pass: persistent wavefronts over
// the list (grid-stride) compute normal, material, G-buffer and the shadow
// ray, so hit pixels run at full lane occupancy and the primary march carries
// none of the shading registers. Same per-ray arithmetic as view_pixel.
struct HitRec {
    uint32_t px, py;
    float t, mat_d;
    float mp[3];
    uint32_t steps, evals;
};

__global__ void hits_reset_kernel(uint32_t *count) { *count = 0u; }

__global__ __launch_bounds__(64) void sparse_march_primary(SparseMarchParams P, HitRec *hits, uint32_t *count) {
    const uint32_t lane = threadIdx.x;
    uint32_t px = blockIdx.x * 8u + (lane & 7u);
    uint32_t py = blockIdx.y * 8u + (lane >> 3);
    bool active = px < P.g.width && py < P.g.height;
    int st = kRayMiss;
    PrimRay r;
    SpRes res;
    if (active) {
        f3 o, dir;
        camera_ray(P.cam, P.g.width, P.g.height, px, py, o, dir);
        prim_init(r, o, dir, P.S.vs, P.su.max_distance);
        if (P.tiles) prim_cull(r, P.tiles[(uint64_t)(py >> 3) * P.tiles_x + (px >> 3)], P.S.vs);
        do {
            st = prim_step<true>(P.S, r, P.su, P.ext.hit_eps_mode, res);
        } while (st == kRayContinue);
        if (st != kRayHit) {
            f3 so;
            (void)prim_finish(P, px, py, r, st, res, so);
        }
    }
    bool hit = active && st == kRayHit;
    uint64_t m = __ballot(hit);
    if (m == 0ull) return;
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(count, (uint32_t)__popcll(m));
    base = (uint32_t)__shfl((int)base, 0);
    if (hit) {
        uint32_t i = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        HitRec h;
        h.px = px;
        h.py = py;
        h.t = res.t;
        h.mat_d = res.mat_d;
        h.mp[0] = res.mat_p.x;
        h.mp[1] = res.mat_p.y;
        h.mp[2] = res.mat_p.z;
        h.steps = r.steps;
        h.evals = r.evals;
        hits[i] = h;
    }
}

__global__ __launch_bounds__(64) void sparse_march_shade(SparseMarchParams P, const HitRec *hits,
                                                         const uint32_t *count) {
    const uint32_t n = *count;
    const float maxd = shadow_maxd(P), k = shadow_k(P);
    const f3 L = mk(P.ext.light_dir[0], P.ext.light_dir[1], P.ext.light_dir[2]);
    for (uint32_t i = blockIdx.x * 64u + threadIdx.x; i < n; i += gridDim.x * 64u) {
        HitRec h = hits[i];
        PrimRay r;
        r.steps = h.steps;
        r.evals = h.evals;
        SpRes res;
        res.t = h.t;
        res.mat_d = h.mat_d;
        res.mat_p = mk(h.mp[0], h.mp[1], h.mp[2]);
        res.hit = true;
        f3 so;
        if (prim_finish(P, h.px, h.py, r, kRayHit, res, so)) {
            ShadowRay sr;
            shadow_init(sr, so, L, P.S.vs);
            while (!shadow_step<true>(P.S, sr, maxd, k)) {
            }
            shadow_store(P, (uint64_t)h.py * P.g.width + h.px, sr);
        }
    }
}

// Tile intervals. The camera basis is inverted once per lane (rows of
// [f r u]^-1 = (r x u, u x f, f x r) / det), so a point's coefficients
// v = a f + b r + c u give its pixel ray coordinates nx = b / a, ny = c / a.
struct TileCam {
    f3 o, ia, ib, ic;
    float w, h, jx, jy;
};

RKF_DEV f3 cross3(f3 a, f3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }

RKF_DEV TileCam tile_cam(const rkf_camera_uniforms &cam, uint32_t w, uint32_t h) {
    const f3 f = mk(cam.forward[0], cam.forward[1], cam.forward[2]);
    const f3 r = mk(cam.right[0], cam.right[1], cam.right[2]);
    const f3 u = mk(cam.up[0], cam.up[1], cam.up[2]);
    const f3 ru = cross3(r, u), uf = cross3(u, f), fr = cross3(f, r);
    const float det = dot(f, ru);
    TileCam t;
    t.o = mk(cam.position[0], cam.position[1], cam.position[2]);
    t.ia = mk(ru.x / det, ru.y / det, ru.z / det);
    t.ib = mk(uf.x / det, uf.y / det, uf.z / det);
    t.ic = mk(fr.x / det, fr.y / det, fr.z / det);
    t.w = (float)w;
    t.h = (float)h;
    t.jx = cam.jitter[0];
    t.jy = cam.jitter[1];
    return t;
}

// One occupied cell: its widened box, the distances from the camera to the
// box (tn) and to its farthest corner (tf), and the rectangle of tiles whose
// rays can meet it. Returns false when no pixel ray can meet it within
// max_distance.
RKF_DEV bool tile_cell(const SparseDev &S, const TileCam &tc, uint32_t cell, float max_distance, uint32_t tiles_x,
                       uint32_t tiles_y, uint32_t &x0, uint32_t &x1, uint32_t &y0, uint32_t &y1, float &tn,
                       float &tf) {
    const uint32_t q[3] = {cell % S.cells[0], (cell / S.cells[0]) % S.cells[1], cell / (S.cells[0] * S.cells[1])};
    const float vs = S.vs;
    const float cw = vs * 64.0f, mg = vs * 9.0f, mb = vs * 9.0f + vs * 16.0f;
    const float o[3] = {tc.o.x, tc.o.y, tc.o.z};
    float lo[3], hi[3];
    float dn2 = 0.0f, df2 = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        lo[a] = S.origin[a] + (float)q[a] * cw - (q[a] == 0u ? mb : mg);
        hi[a] = S.origin[a] + (float)(q[a] + 1u) * cw + (q[a] + 1u == S.cells[a] ? mb : mg);
        const float dl = lo[a] - o[a], dh = o[a] - hi[a];
        const float n = fmaxf(fmaxf(dl, dh), 0.0f);
        const float fa = fmaxf(fabsf(dl), fabsf(dh));
        dn2 = dn2 + n * n;
        df2 = df2 + fa * fa;
    }
    tn = __builtin_sqrtf(dn2);
    tf = __builtin_sqrtf(df2);
    if (tn > max_distance) return false;
    bool all_front = true, any_front = false;
    uint32_t outside = 15u;  // right, left, top, bottom frustum planes
    float nx0 = __builtin_inff(), nx1 = -__builtin_inff(), ny0 = __builtin_inff(), ny1 = -__builtin_inff();
#pragma unroll
    for (uint32_t k = 0; k < 8u; k++) {
        const f3 v = sub(mk((k & 1u) ? hi[0] : lo[0], (k & 2u) ? hi[1] : lo[1], (k & 4u) ? hi[2] : lo[2]), tc.o);
        const float a = dot(tc.ia, v), b = dot(tc.ib, v), c = dot(tc.ic, v);
        const float sl = a + fabsf(a) * 1e-3f;
        if (!(b > sl)) outside &= ~1u;
        if (!(-b > sl)) outside &= ~2u;
        if (!(c > sl)) outside &= ~4u;
        if (!(-c > sl)) outside &= ~8u;
        if (a > 0.0f) {
            any_front = true;
            const float nx = b / a, ny = c / a;
            nx0 = fminf(nx0, nx);
            nx1 = fmaxf(nx1, nx);
            ny0 = fminf(ny0, ny);
            ny1 = fmaxf(ny1, ny);
        } else {
            all_front = false;
        }
    }
    if (outside != 0u) return false;
    if (!any_front && tn > 0.0f) return false;
    if (!all_front) {
        x0 = 0u;
        y0 = 0u;
        x1 = tiles_x - 1u;
        y1 = tiles_y - 1u;
        return true;
    }
    // pixel coordinates of the rectangle (camera_ray inverted), 2 pixels wider
    float fx0 = (nx0 + 1.0f) * 0.5f * tc.w - 0.5f - tc.jx - 2.0f;
    float fx1 = (nx1 + 1.0f) * 0.5f * tc.w - 0.5f - tc.jx + 2.0f;
    float fy0 = (1.0f - ny1) * 0.5f * tc.h - 0.5f - tc.jy - 2.0f;
    float fy1 = (1.0f - ny0) * 0.5f * tc.h - 0.5f - tc.jy + 2.0f;
    fx0 = fminf(fmaxf(fx0, -16.0f), tc.w + 16.0f);
    fx1 = fminf(fmaxf(fx1, -16.0f), tc.w + 16.0f);
    fy0 = fminf(fmaxf(fy0, -16.0f), tc.h + 16.0f);
    fy1 = fminf(fmaxf(fy1, -16.0f), tc.h + 16.0f);
    const int32_t ix0 = (int32_t)floorf(fx0 * 0.125f), ix1 = (int32_t)floorf(fx1 * 0.125f);
    const int32_t iy0 = (int32_t)floorf(fy0 * 0.125f), iy1 = (int32_t)floorf(fy1 * 0.125f);
    if (ix1 < 0 || iy1 < 0 || ix0 > (int32_t)tiles_x - 1 || iy0 > (int32_t)tiles_y - 1) return false;
    x0 = (uint32_t)max(ix0, 0);
    y0 = (uint32_t)max(iy0, 0);
    x1 = (uint32_t)min(ix1, (int32_t)tiles_x - 1);
    y1 = (uint32_t)min(iy1, (int32_t)tiles_y - 1);
    return true;
}

RKF_DEV void tile_merge(uint2 *tiles, uint32_t *cnt, uint64_t i, float tn, float tf) {
    uint32_t *w = reinterpret_cast<uint32_t *>(tiles + i);
    atomicMin(w, __float_as_uint(tn));
    atomicMax(w + 1, __float_as_uint(tf));
    if (cnt) atomicAdd(cnt + i, 1u);
}

__global__ __launch_bounds__(256) void tiles_clear_kernel(uint2 *tiles, uint32_t *cnt, uint32_t *bins, uint32_t n,
                                                          uint32_t *big_count) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i == 0u) *big_count = 0u;
    if (bins && blockIdx.x == 0u) {
        bins[threadIdx.x] = 0u;
        bins[256u + threadIdx.x] = 0u;
    }
    if (i < n) {
        tiles[i] = make_uint2(0x7F800000u, 0u);
        if (cnt) cnt[i] = 0u;
    }
}

// One lane per L1 block; rectangles above kTileInlineMax tiles go to `big`.
__global__ __launch_bounds__(256) void tiles_bin_kernel(SparseDev S, rkf_camera_uniforms cam, float max_distance,
                                                        uint32_t w, uint32_t h, uint2 *tiles, uint32_t *cnt,
                                                        TileBig *big, uint32_t *big_count) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= S.n_blocks) return;
    const uint32_t tiles_x = (w + 7u) / 8u, tiles_y = (h + 7u) / 8u;
    const TileCam tc = tile_cam(cam, w, h);
    uint32_t x0, x1, y0, y1;
    float tn, tf;
    if (!tile_cell(S, tc, S.block_cell[i], max_distance, tiles_x, tiles_y, x0, x1, y0, y1, tn, tf)) return;
    if ((x1 - x0 + 1u) * (y1 - y0 + 1u) <= kTileInlineMax) {
        for (uint32_t y = y0; y <= y1; y++)
            for (uint32_t x = x0; x <= x1; x++) tile_merge(tiles, cnt, (uint64_t)y * tiles_x + x, tn, tf);
        return;
    }
    const uint32_t k = atomicAdd(big_count, 1u);
    big[k] = TileBig{x0, x1, y0, y1, tn, tf};
}

// Large rectangles: one workgroup per record, grid-stride over the records.
__global__ __launch_bounds__(256) void tiles_big_kernel(const TileBig *big, const uint32_t *big_count,
                                                        uint32_t tiles_x, uint2 *tiles, uint32_t *cnt) {
    const uint32_t n = *big_count;
    for (uint32_t k = blockIdx.x; k < n; k += gridDim.x) {
        const TileBig b = big[k];
        const uint32_t rw = b.x1 - b.x0 + 1u, area = rw * (b.y1 - b.y0 + 1u);
        for (uint32_t j = threadIdx.x; j < area; j += 256u)
            tile_merge(tiles, cnt, (uint64_t)(b.y0 + j / rw) * tiles_x + (b.x0 + j % rw), b.tn, b.tf);
    }
}

// Launch order of a culled view's tiles: by the number of occupied cells
// binned into each (capped at 255), most first; within a count, any order.
// A counting sort in two launches: per-workgroup LDS histograms summed into
// `bins[0..255]`, then each workgroup ranks its tiles in LDS, reserves its
// range in every count's slice with one atomic on `bins[256 + count]`, and
// writes them (bins zeroed by tiles_clear_kernel).
__global__ __launch_bounds__(256) void tiles_hist_kernel(const uint32_t *cnt, uint32_t n, uint32_t *bins) {
    __shared__ uint32_t h[kOrderBins];
    const uint32_t tid = threadIdx.x, i = blockIdx.x * 256u + tid;
    h[tid] = 0u;
    __syncthreads();
    if (i < n) atomicAdd(&h[min(cnt[i], kOrderBins - 1u)], 1u);
    __syncthreads();
    if (h[tid]) atomicAdd(&bins[tid], h[tid]);
}

__global__ __launch_bounds__(256) void tiles_order_kernel(const uint32_t *cnt, uint32_t n, uint32_t *bins,
                                                          uint32_t *order) {
    __shared__ uint32_t start[kOrderBins], local[kOrderBins], wsum[4];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6, i = blockIdx.x * 256u + tid;
    // exclusive prefix of the counts' totals in descending count order:
    // thread t holds count 255 - t
    const uint32_t v = bins[kOrderBins - 1u - tid];
    uint32_t incl = v;
#pragma unroll
    for (uint32_t d = 1u; d < 64u; d <<= 1) {
        const uint32_t o = __shfl_up(incl, d, 64);
        if (lane >= d) incl += o;
    }
    if (lane == 63u) wsum[wave] = incl;
    local[tid] = 0u;
    __syncthreads();
    uint32_t before = 0u;
    for (uint32_t w = 0; w < wave; w++) before += wsum[w];
    start[kOrderBins - 1u - tid] = before + incl - v;
    uint32_t bin = 0u, rank = 0u;
    if (i < n) {
        bin = min(cnt[i], kOrderBins - 1u);
        rank = atomicAdd(&local[bin], 1u);
    }
    __syncthreads();
    if (local[tid]) start[tid] += atomicAdd(&bins[kOrderBins + tid], local[tid]);
    __syncthreads();
    if (i < n) order[start[bin] + rank] = i;
}

// Point queries: one lane per point.
__global__ __launch_bounds__(256) void sparse_query_kernel(SparseDev S, const float *xyz, uint32_t n,
                                                           rkfm_point_result *out) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    f3 p = mk(xyz[3 * (uint64_t)i], xyz[3 * (uint64_t)i + 1], xyz[3 * (uint64_t)i + 2]);
    point_query(S, p, out[i]);
}

__global__ __launch_bounds__(256) void sparse_cast_kernel(SparseDev S, const rkfm_sphere_cast *casts, uint32_t n,
                                                          rkfm_cast_result *out) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    sphere_cast(S, casts[i], out[i]);
}

// Contacts: 16 lanes per body (4 bodies per wavefront), one lane per sample
// point; a lane's rank among the body's contacts (deeper first, then lower
// sample index) comes from the 16 lanes' penetrations by shuffles.
__global__ __launch_bounds__(256) void sparse_contacts_kernel(SparseDev S, const rkfm_body *bodies, uint32_t n,
                                                              rkfm_contact *out, uint32_t *counts) {
    const uint32_t gid = blockIdx.x * 256u + threadIdx.x;
    const uint32_t body = gid >> 4, k = gid & 15u;
    const uint32_t lane = threadIdx.x & 63u, base = lane & ~15u;
    bool contact = false;
    float pen = 0.0f;
    f3 wp = mk(0.0f, 0.0f, 0.0f), nrm = mk(0.0f, 0.0f, 0.0f);
    if (body < n) {
        const rkfm_body b = bodies[body];
        f3 lp;
        if (body_sample(b, k, lp)) {
            wp = add(mk(b.position[0], b.position[1], b.position[2]), quat_mul_vec(b.rotation, lp));
            SpCtx c{&S, 0u};
            const float d = sp_sample(c, wp, false, false);
            if (d < b.threshold) {
                contact = true;
                pen = b.threshold - d;
                nrm = field_gradient(S, wp);
            }
        }
    }
    const uint64_t m = __ballot(contact);
    uint32_t rank = 0;
    for (uint32_t j = 0; j < 16u; j++) {
        const float pj = __shfl(pen, (int)(base + j));
        const bool cj = (m >> (base + j)) & 1ull;
        if (cj && (pj > pen || (pj == pen && j < k))) rank++;
    }
    if (body >= n) return;
    if (contact) {
        rkfm_contact r;
        r.position[0] = wp.x;
        r.position[1] = wp.y;
        r.position[2] = wp.z;
        r.normal[0] = nrm.x;
        r.normal[1] = nrm.y;
        r.normal[2] = nrm.z;
        r.penetration = pen;
        r.sample = k;
        out[(uint64_t)body * RKFM_MAX_CONTACTS + rank] = r;
    }
    if (k == 0) counts[body] = (uint32_t)__popcll((m >> base) & 0xFFFFull);
}

}  // namespace rkf

extern "C" hipError_t rkf_launch_sparse_contacts(const rkf::SparseDev *S, const rkfm_body *bodies, uint32_t n,
                                                 rkfm_contact *out, uint32_t *counts, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint64_t threads = (uint64_t)n * 16u;
    hipLaunchKernelGGL(rkf::sparse_contacts_kernel, dim3((unsigned)((threads + 255u) / 256u)), dim3(256), 0, stream, *S,
                       bodies, n, out, counts);
    return hipGetLastError();
}

extern "C" hipError_t rkf_launch_sparse_march(const rkf::SparseMarchParams *P, uint32_t variant, uint32_t waves,
                                              hipStream_t stream) {
    if (variant == 0) {
        dim3 grid((P->g.width + 7u) / 8u, (P->g.height + 7u) / 8u, 1);
        hipLaunchKernelGGL(rkf::sparse_march_plain, grid, dim3(64, 1, 1), 0, stream, *P);
        return hipGetLastError();
    }
    uint32_t tiles = ((P->g.width + 7u) / 8u) * ((P->g.height + 7u) / 8u);
    uint32_t n = waves < tiles ? waves : tiles;
    if (n == 0) n = 1;
    hipLaunchKernelGGL(rkf::sparse_march_tuned, dim3(n), dim3(64), 0, stream, *P, n);
    return hipGetLastError();
}

extern "C" uint64_t rkf_hit_record_bytes() { return sizeof(rkf::HitRec); }

extern "C" hipError_t rkf_launch_sparse_march_split(const rkf::SparseMarchParams *P, void *hits, uint32_t *count,
                                                    uint32_t shade_waves, hipStream_t stream) {
    hipLaunchKernelGGL(rkf::hits_reset_kernel, dim3(1), dim3(1), 0, stream, count);
    dim3 grid((P->g.width + 7u) / 8u, (P->g.height + 7u) / 8u, 1);
    hipLaunchKernelGGL(rkf::sparse_march_primary, grid, dim3(64, 1, 1), 0, stream, *P, (rkf::HitRec *)hits, count);
    uint64_t most = ((uint64_t)P->g.width * P->g.height + 63u) / 64u;
    uint32_t waves = (uint32_t)(most < shade_waves ? most : shade_waves);
    if (waves == 0) waves = 1;
    hipLaunchKernelGGL(rkf::sparse_march_shade, dim3(waves), dim3(64, 1, 1), 0, stream, *P,
                       (const rkf::HitRec *)hits, (const uint32_t *)count);
    return hipGetLastError();
}

extern "C" hipError_t rkf_occupancy_waves(uint32_t which, uint32_t *waves) {
    (void)which;
    int dev = 0, blocks = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, dev);
    if (e != hipSuccess) return e;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, rkf::sparse_march_tuned, 64, 0);
    if (e != hipSuccess) return e;
    *waves = (uint32_t)(blocks > 0 ? blocks : 1) * (uint32_t)prop.multiProcessorCount;
    return hipSuccess;
}

extern "C" hipError_t rkf_launch_sparse_query(const rkf::SparseDev *S, const float *xyz, uint32_t n,
                                              rkfm_point_result *out, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(rkf::sparse_query_kernel, dim3((n + 255u) / 256u, 1, 1), dim3(256, 1, 1), 0, stream, *S, xyz,
                       n, out);
    return hipGetLastError();
}

extern "C" hipError_t rkf_launch_sparse_cast(const rkf::SparseDev *S, const rkfm_sphere_cast *casts, uint32_t n,
                                             rkfm_cast_result *out, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(rkf::sparse_cast_kernel, dim3((n + 255u) / 256u, 1, 1), dim3(256, 1, 1), 0, stream, *S, casts,
                       n, out);
    return hipGetLastError();
}

extern "C" hipError_t rkf_launch_sparse_tiles(const rkf::SparseDev *S, const rkf_camera_uniforms *cam,
                                              float max_distance, uint32_t width, uint32_t height, uint2 *tiles,
                                              uint32_t *cnt, rkf::TileBig *big, uint32_t *big_count, bool clear,
                                              hipStream_t stream) {
    const uint32_t n_tiles = ((width + 7u) / 8u) * ((height + 7u) / 8u);
    // with counts, the 512 words before them are the order's bins; clear =
    // false when the caller knows them cleared (the last march reset them)
    if (clear)
        hipLaunchKernelGGL(rkf::tiles_clear_kernel, dim3((n_tiles + 255u) / 256u), dim3(256), 0, stream, tiles, cnt,
                           cnt ? cnt - 2u * rkf::kOrderBins : nullptr, n_tiles, big_count);
    if (S->n_blocks) {
        hipLaunchKernelGGL(rkf::tiles_bin_kernel, dim3((S->n_blocks + 255u) / 256u), dim3(256), 0, stream, *S, *cam,
                           max_distance, width, height, tiles, cnt, big, big_count);
        hipLaunchKernelGGL(rkf::tiles_big_kernel, dim3(512), dim3(256), 0, stream, (const rkf::TileBig *)big,
                           (const uint32_t *)big_count, (width + 7u) / 8u, tiles, cnt);
    }
    return hipGetLastError();
}

extern "C" hipError_t rkf_launch_sparse_tile_order(uint32_t *cnt, uint32_t n_tiles, uint32_t *order,
                                                   hipStream_t stream) {
    uint32_t *bins = cnt - 2u * rkf::kOrderBins;
    const dim3 grid((n_tiles + 255u) / 256u);
    hipLaunchKernelGGL(rkf::tiles_hist_kernel, grid, dim3(256), 0, stream, (const uint32_t *)cnt, n_tiles, bins);
    hipLaunchKernelGGL(rkf::tiles_order_kernel, grid, dim3(256), 0, stream, (const uint32_t *)cnt, n_tiles, bins,
                       order);
    return hipGetLastError();
}

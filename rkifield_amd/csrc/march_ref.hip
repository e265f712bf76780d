// march_ref.hip — the drop-in ray-march kernel for the reference layouts.
//
// Replaces crates/rkf-render/shaders/ray_march.wgsl:main (dispatched by
// RayMarchPass::dispatch, crates/rkf-render/src/ray_march.rs:96-118): one lane
// per pixel, 8x8 pixels per 64-lane wavefront (= one workgroup), reading the
// AoS VoxelSample pool, dense per-object brick maps, 256-byte GpuObjects,
// per-tile object lists and the coarse field exactly as the reference binds
// them. Control flow follows ray_march.wgsl:874-985 (tiled), :641-773 (BVH),
// :776-858 (brute); normals :1003-1032; G-buffer :1083-1111.
//
// Differences from the WGSL, each deliberate:
//  * coarse field: software trilinear (texel centres, clamp-to-edge) instead
//    of the hardware sampler, so results are reproducible on the CPU;
//  * every buffer index is bounds-checked (out-of-range map entry = EMPTY,
//    out-of-range voxel = word0 0, object index >= num_objects = SDF_TYPE_NONE)
//    where the WGSL relies on robust buffer access;
//  * tile counts above 32 are clamped to 32;
//  * hit epsilon may be relative (hit_threshold * t) when the extension asks.
#include "ref_eval.h"

namespace rkf {

// ray_march.wgsl:228-235, software trilinear. L: lanes per ray (8: the
// ray's lanes share the corner loads; k is the lane's index in the ray).
template <uint32_t L> RKF_DEV float coarse(const RefParams &P, f3 cr, uint32_t k) {
    if (!P.coarse) return 0.0f;
    const rkf_coarse_field_uniforms &u = P.cu;
    f3 fp = sub(cr, mk(u.origin_cam_rel[0], u.origin_cam_rel[1], u.origin_cam_rel[2]));
    float iv = u.inv_voxel_size;
    float ux = fp.x * iv / (float)u.dims[0], uy = fp.y * iv / (float)u.dims[1], uz = fp.z * iv / (float)u.dims[2];
    if (ux < 0.0f || uy < 0.0f || uz < 0.0f || ux > 1.0f || uy > 1.0f || uz > 1.0f) return 0.0f;
    float tc[3] = {ux * (float)u.dims[0] - 0.5f, uy * (float)u.dims[1] - 0.5f, uz * (float)u.dims[2] - 0.5f};
    int32_t i0[3], i1[3];
    float f[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        float fl = floorf(tc[a]);
        f[a] = tc[a] - fl;
        int32_t ii = (int32_t)fl;
        i0[a] = clampi(ii, 0, (int32_t)u.dims[a] - 1);
        i1[a] = clampi(ii + 1, 0, (int32_t)u.dims[a] - 1);
    }
    uint64_t dx = u.dims[0], dxy = (uint64_t)u.dims[0] * u.dims[1];
    auto cf = [&](int32_t x, int32_t y, int32_t z) {
        return h2f(P.coarse[(uint64_t)x + (uint64_t)y * dx + (uint64_t)z * dxy]);
    };
    if (L == 8u) {
        // The ray's 8 lanes load one corner each (lane k: corner (k & 1,
        // k >> 1 & 1, k >> 2)) and reduce the trilinear mix across the
        // group, each level's mix with the sequential code's operand order,
        // so every lane ends with the sequential value.
        const uint32_t kx = k & 1u, ky = (k >> 1) & 1u, kz = k >> 2;
        const float c = cf(kx ? i1[0] : i0[0], ky ? i1[1] : i0[1], kz ? i1[2] : i0[2]);
        const float p1 = dpp_xor1(c);
        const float c1 = kx ? mix(p1, c, f[0]) : mix(c, p1, f[0]);
        const float p2 = dpp_xor2(c1);
        const float c2 = ky ? mix(p2, c1, f[1]) : mix(c1, p2, f[1]);
        const float p3 = dpp_half_mirror(c2);
        return kz ? mix(p3, c2, f[2]) : mix(c2, p3, f[2]);
    }
    float c00 = mix(cf(i0[0], i0[1], i0[2]), cf(i1[0], i0[1], i0[2]), f[0]);
    float c10 = mix(cf(i0[0], i1[1], i0[2]), cf(i1[0], i1[1], i0[2]), f[0]);
    float c01 = mix(cf(i0[0], i0[1], i1[2]), cf(i1[0], i0[1], i1[2]), f[0]);
    float c11 = mix(cf(i0[0], i1[1], i1[2]), cf(i1[0], i1[1], i1[2]), f[0]);
    return mix(mix(c00, c10, f[1]), mix(c01, c11, f[1]), f[2]);
}

struct MarchRes {
    float t;
    bool hit;
    uint32_t mat, obj_id, obj_idx, steps;
};

// ray_march.wgsl:944-962: kBisect halvings of [lo, hi], each keeping the
// half whose start is outside. With 8 lanes per ray (k: the lane's index in
// the ray, from lane `base`) up to three halvings are one round: the 7
// midpoints the next three halvings can take (a binary tree: node n's
// halves are nodes 2n + 1, [lo, mid], and 2n + 2, [mid, hi]) are sampled at
// once, lane k taking node k, each midpoint computed by the additions and
// halvings the one-lane loop makes on that path; the signs then pick the
// path. Evaluations count only the path's samples, as in the loop.
template <uint32_t L>
RKF_DEV void bisect(RefCtx &c, f3 o, f3 sd, float lo, float hi, const rkf_gpu_object &obj, uint32_t oi,
                    uint32_t best_mat, uint32_t best_id, MarchRes &r, uint32_t k, uint32_t base) {
    uint32_t bm = best_mat;
    if (L == 8u) {
        for (uint32_t left = kBisect; left > 0u;) {
            const uint32_t depth = left < 3u ? left : 3u, nodes = (1u << depth) - 1u;
            // this lane's node: the halvings from the root (bit i set: step i keeps [mid, hi])
            uint32_t dirs = 0, dk = 0;
            for (uint32_t n = k; n > 0u && n < nodes; n = (n - 1u) >> 1) {
                dirs = (dirs << 1) | ((n & 1u) ? 0u : 1u);
                dk++;
            }
            float l = lo, h = hi;
            for (uint32_t i = 0; i < dk; i++) {
                const float mid = (l + h) * 0.5f;
                if ((dirs >> i) & 1u) l = mid;
                else h = mid;
            }
            bool neg = false, counted = false;
            uint32_t mm = 0;
            if (k < nodes) {
                const float mid = (l + h) * 0.5f;
                const uint32_t e0 = c.evals;
                neg = eval_obj(c, add(o, muls(sd, mid)), obj, oi, mm) < 0.0f;
                counted = c.evals != e0;
                c.evals = e0;
            }
            const uint32_t negs = (uint32_t)(__ballot(neg) >> base) & 0xFFu;
            const uint32_t cnts = (uint32_t)(__ballot(counted) >> base) & 0xFFu;
            uint32_t n = 0, last_kept = nodes;
            for (uint32_t i = 0; i < depth; i++) {
                const float mid = (lo + hi) * 0.5f;
                c.evals += (cnts >> n) & 1u;
                if ((negs >> n) & 1u) {
                    hi = mid;
                    n = 2u * n + 1u;
                } else {
                    lo = mid;
                    last_kept = n;
                    n = 2u * n + 2u;
                }
            }
            const uint32_t km = lane_read(mm, base + (last_kept < nodes ? last_kept : 0u));
            if (last_kept < nodes) bm = km;
            left -= depth;
        }
        r.t = (lo + hi) * 0.5f;
        r.hit = true;
        r.mat = bm;
        r.obj_id = best_id;
        r.obj_idx = oi;
        return;
    }
    for (uint32_t b = 0; b < kBisect; b++) {
        float mid = (lo + hi) * 0.5f;
        f3 mp = add(o, muls(sd, mid));
        uint32_t mm;
        float md = eval_obj(c, mp, obj, oi, mm);
        if (md < 0.0f) {
            hi = mid;
        } else {
            lo = mid;
            bm = mm;
        }
    }
    r.t = (lo + hi) * 0.5f;
    r.hit = true;
    r.mat = bm;
    r.obj_id = best_id;
    r.obj_idx = oi;
}

constexpr uint32_t kNoObj = 0xFFFFFFFFu;

// The L lanes of one ray each hold the first minimum of their share of the
// step's objects (lane s: objects s, s + L, ...), or (kMaxFloat, kNoObj).
// Combined, every lane holds the first object attaining the minimum over
// all of them, the object ray_march.wgsl:929-938 keeps (strictly smaller
// value, or the same value at a lower index).
// Each exchange is symmetric (distances are finite or kMaxFloat, object
// indices distinct unless both kNoObj), so after the xor-1 and xor-2 levels
// lanes 0-3 and 4-7 each hold one value and the half-row mirror serves as
// the xor-4 level: for 8 lanes the exchanges are DPP moves, not LDS round
// trips.
// G > 1: groups of G lanes already hold one value each (levels below G are
// skipped).
template <uint32_t L, uint32_t G = 1> RKF_DEV void split_min(float &d, uint32_t &m, uint32_t &j) {
    if (L == 8u) {
        auto level = [&](auto xch) {
            const float pd = __uint_as_float(xch(__float_as_uint(d)));
            const uint32_t pm = xch(m), pj = xch(j);
            if (pd < d || (pd == d && pj < j)) {
                d = pd;
                m = pm;
                j = pj;
            }
        };
        if (G <= 1u) level([](uint32_t v) { return dpp_u32<kDppXor1>(v); });
        if (G <= 2u) level([](uint32_t v) { return dpp_u32<kDppXor2>(v); });
        if (G <= 4u) level([](uint32_t v) { return dpp_u32<kDppHalfMirror>(v); });
        return;
    }
#pragma unroll
    for (uint32_t off = 1; off < L; off <<= 1) {
        const float pd = __shfl_xor(d, (int)off);
        const uint32_t pm = (uint32_t)__shfl_xor((int)m, (int)off);
        const uint32_t pj = (uint32_t)__shfl_xor((int)j, (int)off);
        if (pd < d || (pd == d && pj < j)) {
            d = pd;
            m = pm;
            j = pj;
        }
    }
}

// ray_march.wgsl:641-985. Tiled mode: the tile's `count` object records and
// indices, staged in LDS by the workgroup (the workgroup's pixels lie in one
// 16x16 object tile), so a step's object loop reads no global memory. The
// tiled and brute object loops are shared by the ray's L lanes (lane `share`
// takes every L-th object). bis: evaluations of the bisection (made by every
// lane of the ray).
template <uint32_t MODE, uint32_t L, uint32_t G>
RKF_DEV void march_core(RefCtx &c, f3 o, f3 dir, const rkf_gpu_object *tobj, const uint32_t *toi, uint32_t count,
                        uint32_t share, MarchRes &r, uint32_t &bis) {
    const RefParams &P = *c.P;
    r.t = kMaxFloat;
    r.hit = false;
    r.mat = r.obj_id = r.obj_idx = 0;
    r.steps = 0;
    if (MODE == 0) {
        if (count == 0) return;
    } else if (MODE == 1) {
        if (P.su.num_objects == 0) return;
    }
    f3 sd = safe_dir(dir);
    f3 cam = mk(P.cam.position[0], P.cam.position[1], P.cam.position[2]);
    float t = 0.0f, prev_t = 0.0f, prev_dist = kMaxFloat;
    uint32_t near = 0;
    for (uint32_t step = 0; step < kHardMax; step++) {
        if (t > P.su.max_distance || near >= P.su.max_steps) break;
        r.steps++;
        f3 pos = add(o, muls(sd, t));
        if (MODE == 0) {
            float cd = coarse<L>(P, sub(pos, cam), share);
            if (cd > kCoarseNear) {
                prev_dist = kMaxFloat;
                t += cd;
                continue;
            }
        }
        float min_dist = kMaxFloat;
        uint32_t best_mat = 0, best_id = 0, best_idx = 0, best_j = 0;
        if (MODE == 0) {
            // groups of G lanes, each sampling one object per pass (group q:
            // objects q, q + L / G, ...). With one lane per object, a last
            // pass that would hold a single object (9, 17 or 25 objects) is
            // sampled by all the ray's lanes instead, after the minimum over
            // the full passes: that object has the highest index, so it
            // replaces the minimum only when strictly smaller, as in the
            // one-lane loop.
            const bool lone_last = G == 1u && L == 8u && count > L && count % L == 1u;
            const uint32_t q = share / G, sg = share % G, n_split = lone_last ? count - 1u : count;
            float bd = kMaxFloat;
            uint32_t bm = 0, bj = kNoObj;
            for (uint32_t i = q; i < n_split; i += L / G) {
                uint32_t m;
                const float d = eval_obj<G>(c, pos, tobj[i], toi[i], m, sg, threadIdx.x - sg);
                if (d < bd) {
                    bd = d;
                    bm = m;
                    bj = i;
                }
            }
            split_min<L, G>(bd, bm, bj);
            if (G == 1u && L == 8u && lone_last) {
                const uint32_t i = count - 1u;
                uint32_t m;
                const float d = eval_obj<8u>(c, pos, tobj[i], toi[i], m, share, threadIdx.x - share);
                if (d < bd) {
                    bd = d;
                    bm = m;
                    bj = i;
                }
            }
            if (bj != kNoObj) {
                const uint32_t oi = toi[bj];
                min_dist = bd;
                best_mat = bm;
                best_id = oi < P.su.num_objects ? tobj[bj].object_id : 0u;
                best_idx = oi;
                best_j = bj;
            }
        } else if (MODE == 1) {
            uint32_t stack[32];
            uint32_t sp = 1;
            stack[0] = 0;
            while (sp > 0) {
                uint32_t ni = stack[--sp];
                if (ni >= P.n_bvh) continue;
                const rkf_bvh_node &n = P.bvh[ni];
                f3 cl = mk(clampf(pos.x, n.aabb_min[0], n.aabb_max[0]), clampf(pos.y, n.aabb_min[1], n.aabb_max[1]),
                           clampf(pos.z, n.aabb_min[2], n.aabb_max[2]));
                float bd = length(sub(cl, pos));
                if (bd > min_dist) continue;
                if (n.left == RKF_BVH_INVALID) {
                    uint32_t oi = n.right_or_object;
                    if (oi < P.su.num_objects) {
                        uint32_t m;
                        float d = eval_object(c, pos, oi, m);
                        if (d < min_dist) {
                            min_dist = d;
                            best_mat = m;
                            best_id = P.objects[oi].object_id;
                            best_idx = oi;
                        }
                    }
                } else if (sp < 31u) {
                    stack[sp++] = n.left;
                    stack[sp++] = n.right_or_object;
                }
            }
        } else {
            float bd = kMaxFloat;
            uint32_t bm = 0, bj = kNoObj;
            for (uint32_t i = share; i < P.su.num_objects; i += L) {
                uint32_t m;
                const float d = eval_object(c, pos, i, m);
                if (d < bd) {
                    bd = d;
                    bm = m;
                    bj = i;
                }
            }
            split_min<L>(bd, bm, bj);
            if (bj != kNoObj) {
                min_dist = bd;
                best_mat = bm;
                best_id = P.objects[bj].object_id;
                best_idx = bj;
            }
        }
        if (min_dist < 0.0f && prev_dist > hit_eps(P.eps_mode, P.su.hit_threshold, prev_t)) {
            const rkf_gpu_object &bo =
                MODE == 0 ? tobj[best_j] : P.objects[best_idx < P.su.num_objects ? best_idx : 0u];
            const uint32_t e0 = c.evals;
            // the workgroup is one 64-lane wavefront: threadIdx.x is the lane
            bisect<L>(c, o, sd, prev_t, t, bo, best_idx, best_mat, best_id, r, share, threadIdx.x - share);
            bis = c.evals - e0;
            return;
        }
        if (min_dist < hit_eps(P.eps_mode, P.su.hit_threshold, t)) {
            r.t = t;
            r.hit = true;
            r.mat = best_mat;
            r.obj_id = best_id;
            r.obj_idx = best_idx;
            return;
        }
        if (min_dist < kNear) near++;
        prev_t = t;
        prev_dist = min_dist;
        t += fmaxf(min_dist, kMinStep);
    }
}

// march_core for one ray; c.evals then counts the ray's evaluations once
// (the L lanes' shares of the object loops plus one bisection).
template <uint32_t MODE, uint32_t L, uint32_t G = 1>
RKF_DEV void march(RefCtx &c, f3 o, f3 dir, const rkf_gpu_object *tobj, const uint32_t *toi, uint32_t count,
                   uint32_t share, MarchRes &r) {
    uint32_t bis = 0;
    march_core<MODE, L, G>(c, o, dir, tobj, toi, count, share, r, bis);
    if (L > 1) {
        uint32_t e = c.evals - bis;
#pragma unroll
        for (uint32_t off = 1; off < L; off <<= 1) e += (uint32_t)__shfl_xor((int)e, (int)off);
        c.evals = e + bis;
    }
}

// Central-difference tap k of the normal (0 +x, 1 -x, 2 +y, 3 -y, 4 +z,
// 5 -z), the same additions ray_march.wgsl:1008-1013 / :1024-1029 make.
RKF_DEV f3 tap_point(f3 p, float e, uint32_t k) {
    const f3 v = k < 2u ? mk(e, 0.0f, 0.0f) : k < 4u ? mk(0.0f, e, 0.0f) : mk(0.0f, 0.0f, e);
    return (k & 1u) ? sub(p, v) : add(p, v);
}

// The six taps shared by the ray's L lanes (lane `share` takes taps share,
// share + L, ...); every lane receives all six. base: the ray's first lane.
template <uint32_t L, class F> RKF_DEV void shared_taps(uint32_t share, uint32_t base, F &&tap, float tv[6]) {
    constexpr uint32_t Q = (6u + L - 1u) / L;
    float mine[Q];
#pragma unroll
    for (uint32_t q = 0; q < Q; q++) {
        const uint32_t k = share + q * L;
        mine[q] = k < 6u ? tap(k) : 0.0f;
    }
#pragma unroll
    for (uint32_t k = 0; k < 6u; k++) tv[k] = L == 1u ? mine[k] : __shfl(mine[k / L], (int)(base + k % L));
}

// ray_march.wgsl:1003-1032
template <uint32_t L> RKF_DEV f3 normal_for_object(RefCtx &c, f3 pos, uint32_t oi, uint32_t share, uint32_t base) {
    const RefParams &P = *c.P;
    const rkf_gpu_object &o = P.objects[oi];
    float tv[6];
    if (o.sdf_type == RKF_SDF_TYPE_VOXELIZED) {
        f3 lp = xform_point(o.inverse_world, pos);
        float e = o.voxel_size * 2.0f;
        shared_taps<L>(share, base, [&](uint32_t k) { return sample_voxelized(c, tap_point(lp, e, k), o, true); }, tv);
        f3 wg = xform_grad(o.inverse_world, mk(tv[0] - tv[1], tv[2] - tv[3], tv[4] - tv[5]));
        float len = length(wg);
        if (len < 1e-10f) return mk(0.0f, 1.0f, 0.0f);
        return mk(wg.x / len, wg.y / len, wg.z / len);
    }
    float e = P.su.hit_threshold * 10.0f;
    shared_taps<L>(share, base, [&](uint32_t k) {
        uint32_t m;
        return eval_object(c, tap_point(pos, e, k), oi, m);
    }, tv);
    return normalize(mk(tv[0] - tv[1], tv[2] - tv[3], tv[4] - tv[5]));
}

// ray_march.wgsl:1036-1111, one instance per strategy (MODE 0 tiled, 1 BVH,
// 2 brute; :1056-1063 picks it, the host launches that instance). Block = 64
// lanes = 64 / L rays of L lanes each: 8x8 pixels (L 1), 8x4 (2), 4x4 (4),
// 4x2 (8), 2x2 (16). Held to 128 VGPRs (4 waves/SIMD).
template <uint32_t L> struct RefBlock {
    static constexpr uint32_t w = L <= 2u ? 8u : L <= 8u ? 4u : 2u, h = (64u / L) / w;
};
// Workgroups per 16x16 object tile in the ordered tiled launch (8 lanes per
// ray: 4x2-pixel workgroups), and the XCDs the dispatcher deals them to.
constexpr uint32_t kRefTileBlocks = 32u, kXcds = 8u;
template <uint32_t MODE, uint32_t L>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 4))) void ray_march_ref_kernel(RefParams P) {
    static_assert(MODE != 1 || L == 1, "the BVH walk prunes by the running minimum: one lane per ray");
    constexpr uint32_t BW = RefBlock<L>::w, BH = RefBlock<L>::h;
    __shared__ rkf_gpu_object s_obj[RKF_TILE_MAX_OBJECTS];
    __shared__ uint32_t s_oi[RKF_TILE_MAX_OBJECTS];
    const uint32_t lane = threadIdx.x, ray = lane / L, share = lane % L;
    const uint32_t W = P.g.width, H = P.g.height;
    uint32_t bx = blockIdx.x, by = blockIdx.y;
    if (MODE == 0 && P.tile_order) {
        // 1-D launch in the given tile order, (16 / BW) x (16 / BH) = 32
        // workgroups per 16x16 tile. Workgroups go to the XCDs round-robin
        // (workgroup b to XCD b % 8), so each run of 256 workgroups takes 8
        // tiles and gives XCD x all 32 workgroups of the run's tile x: a
        // tile's rays sample its objects' bricks through one L2. Slots past
        // the last tile, and workgroups of edge tiles past the image, leave.
        static_assert(MODE != 0 || (16u % BW == 0u && 16u % BH == 0u && (16u / BW) * (16u / BH) == kRefTileBlocks),
                      "tile blocks");
        const uint32_t ntx = (W + 15u) / 16u;
        const uint32_t run = blockIdx.x / (kRefTileBlocks * kXcds), in_run = blockIdx.x % (kRefTileBlocks * kXcds);
        const uint32_t slot = run * kXcds + in_run % kXcds, sub = in_run / kXcds;
        if (slot >= P.n_tiles) return;
        const uint32_t t = P.tile_order[slot];
        bx = (t % ntx) * (16u / BW) + sub % (16u / BW);
        by = (t / ntx) * (16u / BH) + sub / (16u / BW);
        if (bx * BW >= W || by * BH >= H) return;
    }
    uint32_t px = bx * BW + ray % BW;
    uint32_t py = by * BH + ray / BW;
    uint32_t count = 0;
    if (MODE == 0) {
        // Stage the tile's object list and records (64 lanes, one dword of a
        // 256-byte record each; records past num_objects are never read).
        const uint32_t ntx = (W + 15u) / 16u;
        const uint32_t tile = ((by * BH) / 16u) * ntx + (bx * BW) / 16u;
        count = tile < P.n_tiles ? P.tile_counts[tile] : 0u;
        if (count > RKF_TILE_MAX_OBJECTS) count = RKF_TILE_MAX_OBJECTS;
        if (lane < count) s_oi[lane] = P.tile_idx[(uint64_t)tile * RKF_TILE_MAX_OBJECTS + lane];
        __syncthreads();
        const uint32_t *src = reinterpret_cast<const uint32_t *>(P.objects);
        uint32_t *dst = reinterpret_cast<uint32_t *>(s_obj);
        for (uint32_t j0 = 0; j0 < count; j0 += 8u) {
            uint32_t v[8];
#pragma unroll
            for (uint32_t k = 0; k < 8u; k++) {
                const uint32_t j = j0 + k;
                const uint32_t oi = j < count ? s_oi[j] : ~0u;
                v[k] = oi < P.su.num_objects ? src[(uint64_t)oi * 64u + lane] : 0u;
            }
#pragma unroll
            for (uint32_t k = 0; k < 8u; k++)
                if (j0 + k < count) dst[(j0 + k) * 64u + lane] = v[k];
        }
        __syncthreads();
    }
    if (px >= W || py >= H) return;
    uint64_t p = (uint64_t)py * W + px;
    RefCtx c{&P, 0u};
    f3 o, dir;
    camera_ray(P.cam, W, H, px, py, o, dir);
    MarchRes r;
    if (MODE == 0 && L == 8u) {
        // Lanes per object sample by the tile's object count (uniform in the
        // workgroup): all 8 on one object, 4 on two, 2 on up to four, else 1.
        if (count <= 1u) march<MODE, L, 8>(c, o, dir, s_obj, s_oi, count, share, r);
        else if (count <= 2u) march<MODE, L, 4>(c, o, dir, s_obj, s_oi, count, share, r);
        else if (count <= 4u) march<MODE, L, 2>(c, o, dir, s_obj, s_oi, count, share, r);
        else march<MODE, L, 1>(c, o, dir, s_obj, s_oi, count, share, r);
    } else {
        march<MODE, L>(c, o, dir, s_obj, s_oi, count, share, r);
    }
    // the ray's L lanes hold the same results from here on; the first stores
    if (!r.hit) {
        if (share == 0) {
            store_miss(P.g, p);
            store_debug(P.g, p, mk(0.0f, 0.0f, 0.0f), RKF_EMPTY_SLOT, 0xFFFFFFFFu, 0xFFFFFFFFu, r.steps, c.evals);
        }
        return;
    }
    f3 hp = add(o, muls(dir, r.t));
    uint32_t evals_march = c.evals;
    f3 n = normal_for_object<L>(c, hp, r.obj_idx, share, lane - share);
    const rkf_gpu_object &ho = P.objects[r.obj_idx];
    uint32_t sec = 0, blend = 0, slot = RKF_EMPTY_SLOT, vox = 0xFFFFFFFFu;
    float sdf_at = 0.0f;
    if (ho.sdf_type == RKF_SDF_TYPE_VOXELIZED) {
        f3 lh = xform_point(ho.inverse_world, hp);
        uint64_t i = nearest_voxel(P, lh, ho, slot, vox);
        if (i != ~0ull) {
            sec = (pool_w0(P, i) >> 22) & 0x3Fu;
            blend = (pool_w1(P, i) >> 24) & 0xFFu;
        }
        sdf_at = sample_voxelized(c, lh, ho, false);
    }
    uint32_t packed = (r.mat & 0x3Fu) | (sec << 6) | (blend << 12) | (r.obj_id << 24);
    float mx, my;
    motion_vec(P.cam, W, H, px, py, hp, mx, my);
    if (share != 0) return;
    store_hit(P.g, p, hp, r.t, n, packed, mx, my, sdf_at);
    store_debug(P.g, p, n, slot, vox, r.obj_idx, r.steps, evals_march);
}

}  // namespace rkf

// Lanes per ray of the tiled and brute instances.
constexpr uint32_t kTiledLanes = 8;

extern "C" hipError_t rkf_launch_ref_march(const rkf::RefParams *P, hipStream_t stream) {
    dim3 grid((P->g.width + 7u) / 8u, (P->g.height + 7u) / 8u, 1);
    if (P->mode == 0) {
        using B = rkf::RefBlock<kTiledLanes>;
        dim3 tg((P->g.width + B::w - 1u) / B::w, (P->g.height + B::h - 1u) / B::h, 1);
        if (P->tile_order) tg = dim3((P->n_tiles + rkf::kXcds - 1u) / rkf::kXcds * rkf::kXcds * rkf::kRefTileBlocks, 1, 1);
        if (tg.x == 0u) return hipSuccess;
        hipLaunchKernelGGL((rkf::ray_march_ref_kernel<0, kTiledLanes>), tg, dim3(64, 1, 1), 0, stream, *P);
    } else if (P->mode == 1) {
        hipLaunchKernelGGL((rkf::ray_march_ref_kernel<1, 1>), grid, dim3(64, 1, 1), 0, stream, *P);
    } else {
        using B = rkf::RefBlock<kTiledLanes>;
        dim3 tg((P->g.width + B::w - 1u) / B::w, (P->g.height + B::h - 1u) / B::h, 1);
        hipLaunchKernelGGL((rkf::ray_march_ref_kernel<2, kTiledLanes>), tg, dim3(64, 1, 1), 0, stream, *P);
    }
    return hipGetLastError();
}

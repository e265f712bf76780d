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

// ray_march.wgsl:228-235, software trilinear.
RKF_DEV float coarse(const RefParams &P, f3 cr) {
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

RKF_DEV void bisect(RefCtx &c, f3 o, f3 sd, float lo, float hi, const rkf_gpu_object &obj, uint32_t oi,
                    uint32_t best_mat, uint32_t best_id, MarchRes &r) {
    uint32_t bm = best_mat;
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

// ray_march.wgsl:641-985. Tiled mode: the tile's `count` object records and
// indices, staged in LDS by the workgroup (an 8x8 pixel workgroup lies in one
// 16x16 object tile), so a step's object loop reads no global memory.
template <uint32_t MODE>
RKF_DEV void march(RefCtx &c, f3 o, f3 dir, const rkf_gpu_object *tobj, const uint32_t *toi, uint32_t count,
                   MarchRes &r) {
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
            float cd = coarse(P, sub(pos, cam));
            if (cd > kCoarseNear) {
                prev_dist = kMaxFloat;
                t += cd;
                continue;
            }
        }
        float min_dist = kMaxFloat;
        uint32_t best_mat = 0, best_id = 0, best_idx = 0, best_j = 0;
        if (MODE == 0) {
            for (uint32_t i = 0; i < count; i++) {
                const uint32_t oi = toi[i];
                uint32_t m;
                float d = eval_obj(c, pos, tobj[i], oi, m);
                if (d < min_dist) {
                    min_dist = d;
                    best_mat = m;
                    best_id = oi < P.su.num_objects ? tobj[i].object_id : 0u;
                    best_idx = oi;
                    best_j = i;
                }
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
            for (uint32_t i = 0; i < P.su.num_objects; i++) {
                uint32_t m;
                float d = eval_object(c, pos, i, m);
                if (d < min_dist) {
                    min_dist = d;
                    best_mat = m;
                    best_id = P.objects[i].object_id;
                    best_idx = i;
                }
            }
        }
        if (min_dist < 0.0f && prev_dist > hit_eps(P.eps_mode, P.su.hit_threshold, prev_t)) {
            const rkf_gpu_object &bo =
                MODE == 0 ? tobj[best_j] : P.objects[best_idx < P.su.num_objects ? best_idx : 0u];
            bisect(c, o, sd, prev_t, t, bo, best_idx, best_mat, best_id, r);
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

// ray_march.wgsl:1003-1032
RKF_DEV f3 normal_for_object(RefCtx &c, f3 pos, uint32_t oi) {
    const RefParams &P = *c.P;
    const rkf_gpu_object &o = P.objects[oi];
    if (o.sdf_type == RKF_SDF_TYPE_VOXELIZED) {
        f3 lp = xform_point(o.inverse_world, pos);
        float e = o.voxel_size * 2.0f;
        float gx = sample_voxelized(c, add(lp, mk(e, 0.0f, 0.0f)), o, true) -
                   sample_voxelized(c, sub(lp, mk(e, 0.0f, 0.0f)), o, true);
        float gy = sample_voxelized(c, add(lp, mk(0.0f, e, 0.0f)), o, true) -
                   sample_voxelized(c, sub(lp, mk(0.0f, e, 0.0f)), o, true);
        float gz = sample_voxelized(c, add(lp, mk(0.0f, 0.0f, e)), o, true) -
                   sample_voxelized(c, sub(lp, mk(0.0f, 0.0f, e)), o, true);
        f3 wg = xform_grad(o.inverse_world, mk(gx, gy, gz));
        float len = length(wg);
        if (len < 1e-10f) return mk(0.0f, 1.0f, 0.0f);
        return mk(wg.x / len, wg.y / len, wg.z / len);
    }
    float e = P.su.hit_threshold * 10.0f;
    uint32_t m;
    float nx = eval_object(c, add(pos, mk(e, 0.0f, 0.0f)), oi, m) - eval_object(c, sub(pos, mk(e, 0.0f, 0.0f)), oi, m);
    float ny = eval_object(c, add(pos, mk(0.0f, e, 0.0f)), oi, m) - eval_object(c, sub(pos, mk(0.0f, e, 0.0f)), oi, m);
    float nz = eval_object(c, add(pos, mk(0.0f, 0.0f, e)), oi, m) - eval_object(c, sub(pos, mk(0.0f, 0.0f, e)), oi, m);
    return normalize(mk(nx, ny, nz));
}

// ray_march.wgsl:1036-1111, one instance per strategy (MODE 0 tiled, 1 BVH,
// 2 brute; :1056-1063 picks it, the host launches that instance). Block = 64
// lanes = 8x8 pixels. Held to 128 VGPRs (4 waves/SIMD): a 640x360 view is
// 3600 waves, which then all fit the 256 CUs at once.
template <uint32_t MODE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 4))) void ray_march_ref_kernel(RefParams P) {
    __shared__ rkf_gpu_object s_obj[RKF_TILE_MAX_OBJECTS];
    __shared__ uint32_t s_oi[RKF_TILE_MAX_OBJECTS];
    const uint32_t lane = threadIdx.x;
    uint32_t px = blockIdx.x * 8u + (lane & 7u);
    uint32_t py = blockIdx.y * 8u + (lane >> 3);
    const uint32_t W = P.g.width, H = P.g.height;
    uint32_t count = 0;
    if (MODE == 0) {
        // Stage the tile's object list and records (64 lanes, one dword of a
        // 256-byte record each; records past num_objects are never read).
        const uint32_t ntx = (W + 15u) / 16u;
        const uint32_t tile = ((blockIdx.y * 8u) / 16u) * ntx + (blockIdx.x * 8u) / 16u;
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
    march<MODE>(c, o, dir, s_obj, s_oi, count, r);
    if (!r.hit) {
        store_miss(P.g, p);
        store_debug(P.g, p, mk(0.0f, 0.0f, 0.0f), RKF_EMPTY_SLOT, 0xFFFFFFFFu, 0xFFFFFFFFu, r.steps, c.evals);
        return;
    }
    f3 hp = add(o, muls(dir, r.t));
    uint32_t evals_march = c.evals;
    f3 n = normal_for_object(c, hp, r.obj_idx);
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
    store_hit(P.g, p, hp, r.t, n, packed, mx, my, sdf_at);
    store_debug(P.g, p, n, slot, vox, r.obj_idx, r.steps, evals_march);
}

}  // namespace rkf

extern "C" hipError_t rkf_launch_ref_march(const rkf::RefParams *P, hipStream_t stream) {
    dim3 grid((P->g.width + 7u) / 8u, (P->g.height + 7u) / 8u, 1);
    if (P->mode == 0)
        hipLaunchKernelGGL(rkf::ray_march_ref_kernel<0>, grid, dim3(64, 1, 1), 0, stream, *P);
    else if (P->mode == 1)
        hipLaunchKernelGGL(rkf::ray_march_ref_kernel<1>, grid, dim3(64, 1, 1), 0, stream, *P);
    else
        hipLaunchKernelGGL(rkf::ray_march_ref_kernel<2>, grid, dim3(64, 1, 1), 0, stream, *P);
    return hipGetLastError();
}

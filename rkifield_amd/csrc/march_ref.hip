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
#include "kernels.h"

namespace rkf {

struct RefCtx {
    const RefParams *P;
    uint32_t evals;
};

RKF_DEV uint32_t map_entry(const RefParams &P, uint64_t i) { return i < P.n_maps ? P.maps[i] : RKF_EMPTY_SLOT; }
RKF_DEV uint32_t pool_w0(const RefParams &P, uint64_t i) { return i < P.n_pool_voxels ? P.pool[i].word0 : 0u; }
RKF_DEV uint32_t pool_w1(const RefParams &P, uint64_t i) { return i < P.n_pool_voxels ? P.pool[i].word1 : 0u; }

// ray_march.wgsl:353-413 / :463-498. mat (optional): set to the material
// of the voxel nearest_voxel would read (:540-566) when that voxel is one of
// the 8 corners just loaded (the point lies inside the clamp range, so
// floor(gp / vs) is x0 or x0 + 1 per axis), else left unchanged.
RKF_DEV float sample_voxelized(RefCtx &c, f3 lp, const rkf_gpu_object &o, bool grad, uint32_t *mat = nullptr) {
    float vs = o.voxel_size;
    float be = vs * 8.0f;
    uint32_t dims[3] = {o.brick_map_dims[0], o.brick_map_dims[1], o.brick_map_dims[2]};
    f3 gs = mk((float)dims[0] * be, (float)dims[1] * be, (float)dims[2] * be);
    f3 gp = add(lp, muls(gs, 0.5f));
    float lo = vs * 0.01f;
    f3 cl = mk(clampf(gp.x, lo, gs.x - lo), clampf(gp.y, lo, gs.y - lo), clampf(gp.z, lo, gs.z - lo));
    float outside = length(sub(gp, cl));
    if (outside > be * 2.0f) return outside;
    if (!grad) {
        f3 gmin = mk(o.geometry_aabb_min[0], o.geometry_aabb_min[1], o.geometry_aabb_min[2]);
        f3 gmax = mk(o.geometry_aabb_max[0], o.geometry_aabb_max[1], o.geometry_aabb_max[2]);
        if (gmax.x > gmin.x) {
            f3 gc = mk(clampf(lp.x, gmin.x, gmax.x), clampf(lp.y, gmin.y, gmax.y), clampf(lp.z, gmin.z, gmax.z));
            float gd = length(sub(lp, gc));
            if (gd > be) return gd;
        }
    }
    f3 vc = mk(cl.x / vs - 0.5f, cl.y / vs - 0.5f, cl.z / vs - 0.5f);
    int32_t x0 = (int32_t)floorf(vc.x), y0 = (int32_t)floorf(vc.y), z0 = (int32_t)floorf(vc.z);
    float tx = vc.x - (float)x0, ty = vc.y - (float)y0, tz = vc.z - (float)z0;
    int32_t tot[3] = {(int32_t)dims[0] * 8, (int32_t)dims[1] * 8, (int32_t)dims[2] * 8};
    uint32_t off = o.brick_map_offset;
    const RefParams &P = *c.P;
    if (!grad) c.evals++;
    // sample_voxel_at (:270-288; gradient variant :300-315) for the 8
    // corners: clamped to the grid, EMPTY -> +8 vs, INTERIOR -> -2 vs (-8 vs
    // for gradients), else the pool word's binary16. Two rounds of eight
    // independent loads (map entries, then pool words), so each round waits
    // once. Unused loads read entry 0 (the dispatch keeps both buffers
    // non-empty).
    const int32_t X[2] = {clampi(x0, 0, tot[0] - 1), clampi(x0 + 1, 0, tot[0] - 1)};
    const int32_t Y[2] = {clampi(y0, 0, tot[1] - 1), clampi(y0 + 1, 0, tot[1] - 1)};
    const int32_t Z[2] = {clampi(z0, 0, tot[2] - 1), clampi(z0 + 1, 0, tot[2] - 1)};
    uint32_t ent[8], raw[8];
    bool inm[8], ok[8];
#pragma unroll
    for (uint32_t k = 0; k < 8u; k++) {
        const uint32_t cx = (uint32_t)X[k & 1u], cy = (uint32_t)Y[(k >> 1) & 1u], cz = (uint32_t)Z[k >> 2];
        const uint64_t i = (uint64_t)off + ((cx >> 3) + (cy >> 3) * dims[0] + (cz >> 3) * dims[0] * dims[1]);
        inm[k] = i < P.n_maps;
        ent[k] = P.maps[inm[k] ? i : 0u];
    }
#pragma unroll
    for (uint32_t k = 0; k < 8u; k++) {
        loaded(ent[k]);
        ent[k] = inm[k] ? ent[k] : RKF_EMPTY_SLOT;
    }
#pragma unroll
    for (uint32_t k = 0; k < 8u; k++) {
        const uint32_t cx = (uint32_t)X[k & 1u], cy = (uint32_t)Y[(k >> 1) & 1u], cz = (uint32_t)Z[k >> 2];
        const uint64_t idx = (uint64_t)ent[k] * 512u + (cx & 7u) + (cy & 7u) * 8u + (cz & 7u) * 64u;
        ok[k] = ent[k] < RKF_INTERIOR_SLOT && idx < P.n_pool_voxels;
        raw[k] = P.pool[ok[k] ? idx : 0u].word0;
    }
    float cv[8];
#pragma unroll
    for (uint32_t k = 0; k < 8u; k++) {
        loaded(raw[k]);
        // pool_w0: words past the pool read 0
        cv[k] = ent[k] == RKF_EMPTY_SLOT      ? vs * 8.0f
                : ent[k] == RKF_INTERIOR_SLOT ? (grad ? -(vs * 8.0f) : -(vs * 2.0f))
                                              : h2f((uint16_t)(ok[k] ? raw[k] : 0u));
    }
    if (mat && cl.x == gp.x && cl.y == gp.y && cl.z == gp.z) {
        // nearest_voxel: floor(gp / vs) clamped to the grid, EMPTY / INTERIOR -> 0
        const uint32_t k = ((int32_t)floorf(cl.x / vs) != x0 ? 1u : 0u) |
                           ((int32_t)floorf(cl.y / vs) != y0 ? 2u : 0u) |
                           ((int32_t)floorf(cl.z / vs) != z0 ? 4u : 0u);
        uint32_t e = ent[0], w = raw[0];
#pragma unroll
        for (uint32_t j = 1; j < 8u; j++) {
            e = k == j ? ent[j] : e;
            w = k == j ? (ok[j] ? raw[j] : 0u) : w;
        }
        if (k == 0u) w = ok[0] ? raw[0] : 0u;
        *mat = e >= RKF_INTERIOR_SLOT ? 0u : ((w >> 16) & 0x3Fu);
    }
    const float c000 = cv[0], c100 = cv[1], c010 = cv[2], c110 = cv[3];
    const float c001 = cv[4], c101 = cv[5], c011 = cv[6], c111 = cv[7];
    float c00 = mix(c000, c100, tx), c10 = mix(c010, c110, tx);
    float c01 = mix(c001, c101, tx), c11 = mix(c011, c111, tx);
    float c0 = mix(c00, c10, ty), c1 = mix(c01, c11, ty);
    return mix(c0, c1, tz) + outside;
}

// Shared nearest-voxel lookup of ray_march.wgsl:540-566 / :570-596.
RKF_DEV uint64_t nearest_voxel(const RefParams &P, f3 lp, const rkf_gpu_object &o, uint32_t &slot_out,
                               uint32_t &vox_out) {
    float vs = o.voxel_size;
    float be = vs * 8.0f;
    uint32_t d0 = o.brick_map_dims[0], d1 = o.brick_map_dims[1], d2 = o.brick_map_dims[2];
    f3 gs = mk((float)d0 * be, (float)d1 * be, (float)d2 * be);
    f3 gp = add(lp, muls(gs, 0.5f));
    slot_out = RKF_EMPTY_SLOT;
    vox_out = 0xFFFFFFFFu;
    if (gp.x < 0.0f || gp.y < 0.0f || gp.z < 0.0f || gp.x >= gs.x || gp.y >= gs.y || gp.z >= gs.z) return ~0ull;
    int32_t cx = clampi((int32_t)floorf(gp.x / vs), 0, (int32_t)d0 * 8 - 1);
    int32_t cy = clampi((int32_t)floorf(gp.y / vs), 0, (int32_t)d1 * 8 - 1);
    int32_t cz = clampi((int32_t)floorf(gp.z / vs), 0, (int32_t)d2 * 8 - 1);
    uint32_t flat = (uint32_t)(cx / 8) + (uint32_t)(cy / 8) * d0 + (uint32_t)(cz / 8) * d0 * d1;
    uint32_t slot = map_entry(P, (uint64_t)o.brick_map_offset + flat);
    slot_out = slot;
    uint32_t lv = (uint32_t)(cx % 8) + (uint32_t)(cy % 8) * 8u + (uint32_t)(cz % 8) * 64u;
    if (slot == RKF_EMPTY_SLOT || slot == RKF_INTERIOR_SLOT) return ~0ull;
    vox_out = lv;
    return (uint64_t)slot * 512u + lv;
}

// ray_march.wgsl:182-264
RKF_DEV float analytical(f3 p, const rkf_gpu_object &o) {
    float p0 = o.sdf_params[0], p1 = o.sdf_params[1], p2 = o.sdf_params[2];
    switch (o.primitive_type) {
    case 0: return length(p) - p0;
    case 1: {
        f3 q = mk(fabsf(p.x) - p0, fabsf(p.y) - p1, fabsf(p.z) - p2);
        return length(mk(fmaxf(q.x, 0.0f), fmaxf(q.y, 0.0f), fmaxf(q.z, 0.0f))) +
               fminf(fmaxf(q.x, fmaxf(q.y, q.z)), 0.0f);
    }
    case 2: return length(mk(p.x, fmaxf(fabsf(p.y) - p1, 0.0f), p.z)) - p0;
    case 3: {
        float qx = __builtin_sqrtf(p.x * p.x + p.z * p.z) - p0;
        return __builtin_sqrtf(qx * qx + p.y * p.y) - p1;
    }
    case 4: {
        float dx = __builtin_sqrtf(p.x * p.x + p.z * p.z) - p0, dy = fabsf(p.y) - p1;
        float mx = fmaxf(dx, 0.0f), my = fmaxf(dy, 0.0f);
        return fminf(fmaxf(dx, dy), 0.0f) + __builtin_sqrtf(mx * mx + my * my);
    }
    case 5: {
        f3 n = normalize(mk(p0, p1, p2));
        return dot(p, n) + 0.0f;
    }
    default: return kMaxFloat;
    }
}

// ray_march.wgsl:600-634, for object record o (index oi).
RKF_DEV float eval_obj(RefCtx &c, f3 wp, const rkf_gpu_object &o, uint32_t oi, uint32_t &mat) {
    const RefParams &P = *c.P;
    mat = 0;
    if (oi >= P.su.num_objects) return kMaxFloat;
    uint32_t st = o.sdf_type;
    if (st == RKF_SDF_TYPE_NONE) return kMaxFloat;
    f3 lp = xform_point(o.inverse_world, wp);
    float d;
    if (st == RKF_SDF_TYPE_ANALYTICAL) {
        d = analytical(lp, o);
        // evaluate_object hands the material on as an f32 (:633); the march
        // converts it back with u32() (saturating)
        const float mf = (float)o.material_id;
        mat = mf >= 4294967295.0f ? 0xFFFFFFFFu : (uint32_t)mf;
    } else {
        uint32_t near_mat = ~0u;
        d = sample_voxelized(c, lp, o, false, &near_mat);
        if (d < o.voxel_size * 16.0f) {
            if (near_mat != ~0u) {
                mat = near_mat;  // the nearest voxel was a loaded corner
            } else {
                uint32_t sl, vx;
                uint64_t i = nearest_voxel(P, lp, o, sl, vx);
                mat = i == ~0ull ? 0u : ((pool_w0(P, i) >> 16) & 0x3Fu);
            }
        }
    }
    float ms = fminf(o.accumulated_scale[0], fminf(o.accumulated_scale[1], o.accumulated_scale[2]));
    return d * ms;
}

RKF_DEV float eval_object(RefCtx &c, f3 wp, uint32_t oi, uint32_t &mat) {
    const RefParams &P = *c.P;
    return eval_obj(c, wp, P.objects[oi < P.su.num_objects ? oi : 0u], oi, mat);
}

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

// ref_eval.h — evaluate_object over the reference layouts (AoS VoxelSample
// pool, dense per-object brick maps, 256-byte GpuObjects): the field the
// drop-in march samples (crates/rkf-render/shaders/ray_march.wgsl:182-264,
// :270-413, :463-634). Shared by the drop-in kernel (march_ref.hip) and the
// reference-scene to sparse-scene converter (sparse_build.hip).
#pragma once

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
//
// G lanes share one sample (G = 1, 2, 4, 8; every lane of the group passes
// the same point and object and must be active): lane s (0 <= s < G, the
// group's first lane is gbase) loads corners k = s + G * j for j < 8 / G
// (corner k = (k & 1, k >> 1 & 1, k >> 2)), the group mixes across its
// lanes by DPP exchanges at the levels its lanes span (x for G >= 2, y for
// G >= 4, z for G = 8), each mix with the one-lane operand order, so every
// lane returns the one-lane value. Only lane 0 counts the evaluation.
template <uint32_t G = 1>
RKF_DEV float sample_voxelized(RefCtx &c, f3 lp, const rkf_gpu_object &o, bool grad, uint32_t *mat = nullptr,
                               uint32_t s = 0, uint32_t gbase = 0) {
    static_assert(G == 1 || G == 2 || G == 4 || G == 8, "corner groups of 1, 2, 4 or 8 lanes");
    constexpr uint32_t NC = 8u / G;  // corners per lane
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
    if (!grad && s == 0u) c.evals++;
    // sample_voxel_at (:270-288; gradient variant :300-315) for the
    // corners: clamped to the grid, EMPTY -> +8 vs, INTERIOR -> -2 vs (-8 vs
    // for gradients), else the pool word's binary16. Two rounds of
    // independent loads (map entries, then pool words), so each round waits
    // once. Unused loads read entry 0 (the dispatch keeps both buffers
    // non-empty).
    const int32_t X[2] = {clampi(x0, 0, tot[0] - 1), clampi(x0 + 1, 0, tot[0] - 1)};
    const int32_t Y[2] = {clampi(y0, 0, tot[1] - 1), clampi(y0 + 1, 0, tot[1] - 1)};
    const int32_t Z[2] = {clampi(z0, 0, tot[2] - 1), clampi(z0 + 1, 0, tot[2] - 1)};
    // Indices in 32 bits: the uploads hold n_maps and n_pool_voxels below
    // 2^32, so an offset + flat sum that carries, or a slot of 2^23 or more,
    // lies past its buffer exactly when the 64-bit index would.
    const uint32_t n_maps = (uint32_t)P.n_maps, n_pool = (uint32_t)P.n_pool_voxels;
    uint32_t ent[NC], raw[NC], cx[NC], cy[NC], cz[NC];
    bool inm[NC], ok[NC];
#pragma unroll
    for (uint32_t j = 0; j < NC; j++) {
        const uint32_t k = s + G * j;
        cx[j] = (uint32_t)X[k & 1u];
        cy[j] = (uint32_t)Y[(k >> 1) & 1u];
        cz[j] = (uint32_t)Z[k >> 2];
        const uint32_t i = off + ((cx[j] >> 3) + (cy[j] >> 3) * dims[0] + (cz[j] >> 3) * dims[0] * dims[1]);
        inm[j] = i >= off && i < n_maps;
        ent[j] = P.maps[inm[j] ? i : 0u];
    }
#pragma unroll
    for (uint32_t j = 0; j < NC; j++) {
        loaded(ent[j]);
        ent[j] = inm[j] ? ent[j] : RKF_EMPTY_SLOT;
    }
#pragma unroll
    for (uint32_t j = 0; j < NC; j++) {
        const uint32_t idx = (ent[j] << 9) | ((cx[j] & 7u) + (cy[j] & 7u) * 8u + (cz[j] & 7u) * 64u);
        ok[j] = ent[j] < (1u << 23) && idx < n_pool;
        raw[j] = P.pool[ok[j] ? idx : 0u].word0;
    }
    float cv[NC];
    uint32_t mats = 0;  // corner j's material in byte j (G > 1)
#pragma unroll
    for (uint32_t j = 0; j < NC; j++) {
        loaded(raw[j]);
        // pool_w0: words past the pool read 0
        const uint32_t w = ok[j] ? raw[j] : 0u;
        cv[j] = ent[j] == RKF_EMPTY_SLOT      ? vs * 8.0f
                : ent[j] == RKF_INTERIOR_SLOT ? (grad ? -(vs * 8.0f) : -(vs * 2.0f))
                                              : h2f((uint16_t)w);
        if (G > 1u) mats |= (ent[j] >= RKF_INTERIOR_SLOT ? 0u : ((w >> 16) & 0x3Fu)) << (8u * j);
    }
    if (mat && cl.x == gp.x && cl.y == gp.y && cl.z == gp.z) {
        // nearest_voxel: floor(gp / vs) clamped to the grid, EMPTY / INTERIOR -> 0
        const uint32_t k = ((int32_t)floorf(cl.x / vs) != x0 ? 1u : 0u) |
                           ((int32_t)floorf(cl.y / vs) != y0 ? 2u : 0u) |
                           ((int32_t)floorf(cl.z / vs) != z0 ? 4u : 0u);
        if (G == 1u) {
            uint32_t e = ent[0], w = raw[0];
#pragma unroll
            for (uint32_t j = 1; j < NC; j++) {
                e = k == j ? ent[j] : e;
                w = k == j ? (ok[j] ? raw[j] : 0u) : w;
            }
            if (k == 0u) w = ok[0] ? raw[0] : 0u;
            *mat = e >= RKF_INTERIOR_SLOT ? 0u : ((w >> 16) & 0x3Fu);
        } else {
            // corner k sits in lane k % G, byte k / G (the whole group takes this branch)
            *mat = (lane_read(mats, gbase + k % G) >> (8u * (k / G))) & 0xFFu;
        }
    }
    float r;
    if (G == 1u) {
        const float c000 = cv[0], c100 = cv[1 % NC], c010 = cv[2 % NC], c110 = cv[3 % NC];
        const float c001 = cv[4 % NC], c101 = cv[5 % NC], c011 = cv[6 % NC], c111 = cv[7 % NC];
        float c00 = mix(c000, c100, tx), c10 = mix(c010, c110, tx);
        float c01 = mix(c001, c101, tx), c11 = mix(c011, c111, tx);
        float c0 = mix(c00, c10, ty), c1 = mix(c01, c11, ty);
        r = mix(c0, c1, tz);
    } else {
        const uint32_t kx = s & 1u, ky = (s >> 1) & 1u;
        // x level (lanes s, s ^ 1): cv[j] becomes c_yz of its (y, z)
#pragma unroll
        for (uint32_t j = 0; j < NC; j++) {
            const float p = dpp_xor1(cv[j]);
            cv[j] = kx ? mix(p, cv[j], tx) : mix(cv[j], p, tx);
        }
        float c0, c1;  // the z = 0 and z = 1 values of the y level
        if (G == 2u) {  // j = y + 2 z
            c0 = mix(cv[0], cv[1 % NC], ty);
            c1 = mix(cv[2 % NC], cv[3 % NC], ty);
        } else {  // y level across lanes s, s ^ 2; j = z (G = 4) or none (G = 8)
#pragma unroll
            for (uint32_t j = 0; j < NC; j++) {
                const float p = dpp_xor2(cv[j]);
                cv[j] = ky ? mix(p, cv[j], ty) : mix(cv[j], p, ty);
            }
            c0 = cv[0];
            c1 = cv[1 % NC];
        }
        if (G == 8u) {  // z level across the two halves (lanes 0-3 and 4-7 each hold one value)
            const float p = dpp_half_mirror(c0);
            r = (s >> 2) ? mix(p, c0, tz) : mix(c0, p, tz);
        } else {
            r = mix(c0, c1, tz);
        }
    }
    return r + outside;
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

// ray_march.wgsl:600-634, for object record o (index oi); voxelized
// objects sampled by a group of G lanes as sample_voxelized describes.
template <uint32_t G = 1>
RKF_DEV float eval_obj(RefCtx &c, f3 wp, const rkf_gpu_object &o, uint32_t oi, uint32_t &mat, uint32_t s = 0,
                       uint32_t gbase = 0) {
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
        d = sample_voxelized<G>(c, lp, o, false, &near_mat, s, gbase);
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

}  // namespace rkf

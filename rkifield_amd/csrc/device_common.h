// device_common.h — gfx950 device helpers shared by the march kernels.
//
// Numerics contract: every file is compiled with -ffp-contract=off (no FMA
// contraction) and HIP's default correctly rounded fp32 division and sqrt, so
// each float operation below rounds exactly once, in the order written. The
// CPU twin in oracle/trace_impl.h uses the same order.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#include "../../include/rkf_march.h"

#define RKF_DEV __device__ __forceinline__

namespace rkf {

constexpr float kMaxFloat = 3.402823e+38f;  // ray_march.wgsl:28
constexpr float kMinStep = 0.0005f;         // ray_march.wgsl:30
constexpr uint32_t kBisect = 8u;            // ray_march.wgsl:31
constexpr uint32_t kHardMax = 1024u;        // ray_march.wgsl:36
constexpr float kNear = 0.02f;              // ray_march.wgsl:41
constexpr float kCoarseNear = 0.5f;         // ray_march.wgsl:863

struct f3 {
    float x, y, z;
};
// Marks a loaded value as used here: every load before it is issued before
// any code after it, so a group of independent loads waits once instead of
// being sunk one by one into the branches that consume them.
RKF_DEV void loaded(uint32_t &v) { asm volatile("" : "+v"(v)); }

// A pointer loaded from memory (not a kernel argument) that points to global
// memory, re-typed so its loads and stores compile to global_* instead of
// flat_* instructions (flat ones also count against lgkmcnt and check the
// address aperture). The empty asm keeps the compiler from folding the
// address-space round trip away; the value stays in SGPRs.
template <class T> RKF_DEV T *global_ptr(T *p) {
    __attribute__((address_space(1))) T *g = (__attribute__((address_space(1))) T *)p;
    asm("" : "+s"(g));
    return (T *)g;
}

RKF_DEV f3 mk(float x, float y, float z) { return f3{x, y, z}; }
RKF_DEV f3 add(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
RKF_DEV f3 sub(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
RKF_DEV f3 muls(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
RKF_DEV float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RKF_DEV float length(f3 a) { return __builtin_sqrtf(dot(a, a)); }
RKF_DEV float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
RKF_DEV int32_t clampi(int32_t x, int32_t lo, int32_t hi) { return x < lo ? lo : (x > hi ? hi : x); }
// WGSL mix: e1 * (1 - e3) + e2 * e3.
RKF_DEV float mix(float a, float b, float t) { return a * (1.0f - t) + b * t; }

// DPP lane exchanges inside aligned 8-lane groups: xor 1 and xor 2 (quad
// permutations) and the half-row mirror (lane i <-> 7 - i), which equals
// xor 4 once lanes 0-3 and lanes 4-7 each hold one value. Every lane of the
// group must be active.
template <int kCtrl> RKF_DEV uint32_t dpp_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, kCtrl, 0xF, 0xF, false);
}
constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppHalfMirror = 0x141;
RKF_DEV float dpp_xor1(float v) { return __uint_as_float(dpp_u32<kDppXor1>(__float_as_uint(v))); }
RKF_DEV float dpp_xor2(float v) { return __uint_as_float(dpp_u32<kDppXor2>(__float_as_uint(v))); }
RKF_DEV float dpp_half_mirror(float v) { return __uint_as_float(dpp_u32<kDppHalfMirror>(__float_as_uint(v))); }
// Value of v in lane src of the wavefront (src must be active).
RKF_DEV uint32_t lane_read(uint32_t v, uint32_t src) {
    return (uint32_t)__builtin_amdgcn_ds_bpermute((int)(src << 2), (int)v);
}
RKF_DEV f3 normalize(f3 a) {
    float l = length(a);
    return f3{a.x / l, a.y / l, a.z / l};
}
RKF_DEV f3 safe_dir(f3 d) {
    return f3{fabsf(d.x) < 1e-10f ? 1e-10f : d.x, fabsf(d.y) < 1e-10f ? 1e-10f : d.y,
              fabsf(d.z) < 1e-10f ? 1e-10f : d.z};
}

// binary16 <-> binary32 (hardware v_cvt: exact widening, RNE narrowing).
RKF_DEV float h2f(uint32_t bits16) {
    return __half2float(__ushort_as_half((unsigned short)(bits16 & 0xFFFFu)));
}
RKF_DEV uint32_t f2h(float f) { return (uint32_t)__half_as_ushort(__float2half_rn(f)); }

// inverse_world * vec4(p, 1), column-major m[c*4+r], summed left to right.
RKF_DEV f3 xform_point(const float *m, f3 p) {
    return f3{m[0] * p.x + m[4] * p.y + m[8] * p.z + m[12], m[1] * p.x + m[5] * p.y + m[9] * p.z + m[13],
              m[2] * p.x + m[6] * p.y + m[10] * p.z + m[14]};
}
// transpose(inverse_world) * vec4(g, 0) (ray_march.wgsl:1014).
RKF_DEV f3 xform_grad(const float *m, f3 g) {
    return f3{m[0] * g.x + m[1] * g.y + m[2] * g.z + m[3] * 0.0f, m[4] * g.x + m[5] * g.y + m[6] * g.z + m[7] * 0.0f,
              m[8] * g.x + m[9] * g.y + m[10] * g.z + m[11] * 0.0f};
}

RKF_DEV float hit_eps(uint32_t mode, float thr, float t) { return mode == RKFM_EPS_RELATIVE ? thr * t : thr; }

// Ray generation (ray_march.wgsl:1046-1053).
RKF_DEV void camera_ray(const rkf_camera_uniforms &cam, uint32_t w, uint32_t h, uint32_t px, uint32_t py, f3 &o,
                        f3 &dir) {
    float ux = ((float)px + 0.5f + cam.jitter[0]) / (float)w;
    float uy = ((float)py + 0.5f + cam.jitter[1]) / (float)h;
    float nx = ux * 2.0f - 1.0f, ny = 1.0f - uy * 2.0f;
    o = f3{cam.position[0], cam.position[1], cam.position[2]};
    f3 f{cam.forward[0], cam.forward[1], cam.forward[2]};
    f3 r{cam.right[0], cam.right[1], cam.right[2]};
    f3 u{cam.up[0], cam.up[1], cam.up[2]};
    dir = normalize(add(add(f, muls(r, nx)), muls(u, ny)));
}

// Motion vector (ray_march.wgsl:1096-1103).
RKF_DEV void motion_vec(const rkf_camera_uniforms &cam, uint32_t w, uint32_t h, uint32_t px, uint32_t py, f3 hp,
                        float &mx, float &my) {
    const float *m = cam.prev_vp;
    float cx = m[0] * hp.x + m[4] * hp.y + m[8] * hp.z + m[12];
    float cy = m[1] * hp.x + m[5] * hp.y + m[9] * hp.z + m[13];
    float cw = m[3] * hp.x + m[7] * hp.y + m[11] * hp.z + m[15];
    mx = 0.0f;
    my = 0.0f;
    if (cw > 0.0f) {
        float nx = cx / cw, ny = cy / cw;
        float pu = nx * 0.5f + 0.5f, pv = 0.5f - ny * 0.5f;
        float cu = ((float)px + 0.5f) / (float)w, cv = ((float)py + 0.5f) / (float)h;
        mx = cu - pu;
        my = cv - pv;
    }
}

struct GbufPtrs {
    float4 *position;
    uint2 *normal;
    uint32_t *material;
    float4 *motion;
    rkfm_debug_px *debug;
    float *shadow;
    uint32_t width, height;
};

// G-buffer write (ray_march.wgsl:1083-1111).
RKF_DEV void store_hit(const GbufPtrs &g, uint64_t p, f3 hp, float t, f3 n, uint32_t packed, float mx, float my,
                       float sdf_at) {
    g.position[p] = make_float4(hp.x, hp.y, hp.z, t);
    uint32_t hx = f2h(n.x), hy = f2h(n.y), hz = f2h(n.z);
    g.normal[p] = make_uint2(hx | (hy << 16), hz);
    g.material[p] = packed;
    g.motion[p] = make_float4(mx, my, 1.0f, sdf_at);
}
RKF_DEV void store_miss(const GbufPtrs &g, uint64_t p) {
    g.position[p] = make_float4(0.0f, 0.0f, 0.0f, kMaxFloat);
    g.normal[p] = make_uint2(0u, 0u);
    g.material[p] = 0u;
    g.motion[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}
RKF_DEV void store_debug(const GbufPtrs &g, uint64_t p, f3 n, uint32_t slot, uint32_t vox, uint32_t cell,
                         uint32_t steps, uint32_t evals, uint32_t sh_steps = 0, uint32_t sh_evals = 0) {
    if (!g.debug) return;
    rkfm_debug_px d;
    d.normal[0] = n.x;
    d.normal[1] = n.y;
    d.normal[2] = n.z;
    d.brick_slot = slot;
    d.voxel_index = vox;
    d.cell_index = cell;
    d.steps = steps;
    d.evals = evals;
    d.shadow_steps = sh_steps;
    d.shadow_evals = sh_evals;
    d._pad[0] = 0;
    d._pad[1] = 0;
    g.debug[p] = d;
}

}  // namespace rkf

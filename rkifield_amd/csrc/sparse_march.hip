// sparse_march.hip — single-view kernels over a resident sparse scene and the
// point-query kernel. The per-ray arithmetic lives in sparse_device.h.
#include "sparse_device.h"

namespace rkf {

// Plain variant: one lane per pixel, 8x8 pixels per wavefront, each ray
// marched to completion, then its shadow ray.
__global__ __launch_bounds__(64) void sparse_march_plain(SparseMarchParams P) {
    uint32_t px = blockIdx.x * 8u + (threadIdx.x & 7u);
    uint32_t py = blockIdx.y * 8u + (threadIdx.x >> 3);
    if (px >= P.g.width || py >= P.g.height) return;
    view_pixel(P, px, py);
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
                        prim_init(r, o, dir, P.S.vs);
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
                int st = prim_step(P.S, r, P.su, P.ext.hit_eps_mode, res);
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
                if (shadow_step(P.S, r, maxd, ksh)) {
                    shadow_store(P, (uint64_t)py * W + px, r);
                    kind = 0;
                }
            }
        }
    }
}

// Point queries: one lane per point.
__global__ __launch_bounds__(256) void sparse_query_kernel(SparseDev S, const float *xyz, uint32_t n,
                                                           rkfm_point_result *out) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    f3 p = mk(xyz[3 * (uint64_t)i], xyz[3 * (uint64_t)i + 1], xyz[3 * (uint64_t)i + 2]);
    point_query(S, p, out[i]);
}

}  // namespace rkf

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

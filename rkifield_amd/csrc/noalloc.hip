// noalloc.hip — the sparse march for callers that cannot allocate device
// memory (code under HIP graph capture, serving loops, harnesses that forbid
// hipMalloc).
//
// rkfm_sparse_march keeps the scratch of a culled march (per-tile ray
// intervals, the list of cells whose tile rectangle is too large for one
// lane, and that list's counter) in context buffers that grow on demand.
// rkfm_sparse_march_noalloc runs the same kernels on the same inputs with
// that scratch in `scratch`, one fixed array per device that the library
// holds until the process ends. Host code never uses the array's name as an
// address: rkf_noalloc_init (called by rkfm_create with the device current)
// looks its device address up once per device, and the entry point passes
// that address to the kernels.
#include <cstring>
#include <mutex>

#include "host_internal.h"

namespace {

// Layout of the scratch array: the large-rectangle counter, the per-tile
// (t_lo, t_hi) pairs, the large-rectangle records.
constexpr uint64_t kCountBytes = 256;
constexpr uint64_t kTilesBytes = (uint64_t)RKFM_NOALLOC_MAX_TILES * 8u;
constexpr uint64_t kBigOffset = kCountBytes + kTilesBytes;
static_assert(kBigOffset + (uint64_t)RKFM_NOALLOC_MAX_BLOCKS * sizeof(rkf::TileBig) <= RKFM_NOALLOC_SCRATCH_BYTES,
              "scratch layout exceeds the scratch array");

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};

}  // namespace

static __device__ __align__(256) unsigned char scratch[RKFM_NOALLOC_SCRATCH_BYTES];

// Once per device, with that device current: the device address of this
// device's copy of `scratch`.
int rkf_noalloc_init(int device) {
    if (device < 0 || device >= kMaxDevices) return RKFM_OK;  // the entry point reports it
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device]) return RKFM_OK;
    void *p = nullptr;
    int rc = rkf_hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "scratch symbol address");
    if (!rc) g_ptr[device] = p;
    return rc;
}

extern "C" int rkfm_sparse_march_noalloc(rkfm_ctx *c, const rkfm_sparse_scene *s, const rkf_camera_uniforms *cam,
                                         const rkf_scene_uniforms *su, const rkfm_scene_ext *ext,
                                         const rkfm_gbuffer *out, uint32_t variant, void *stream) {
    if (!c || !s || !cam || !su || !out) return rkf_set_error(RKFM_ERR_INVALID_ARG, "march_noalloc: null argument");
    if (variant > 1) return rkf_set_error(RKFM_ERR_INVALID_ARG, "march_noalloc: variant %u (0 or 1)", variant);
    if (ext && ext->hit_eps_mode > RKFM_EPS_RELATIVE)
        return rkf_set_error(RKFM_ERR_INVALID_ARG, "march_noalloc: hit_eps_mode %u", ext->hit_eps_mode);
    if (ext && (ext->march_flags & ~RKFM_MARCH_TILE_CULL))
        return rkf_set_error(RKFM_ERR_INVALID_ARG, "march_noalloc: march_flags 0x%x", ext->march_flags);
    void *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (c->device >= 0 && c->device < kMaxDevices) base = g_ptr[c->device];
    }
    if (!base) return rkf_set_error(RKFM_ERR_NOT_READY, "march_noalloc: no scratch on device %d", c->device);
    rkf::SparseMarchParams P;
    memset(&P, 0, sizeof(P));
    int rc = rkf_fill_gbuf(out, P.g);
    if (rc) return rc;
    P.S = s->dev();
    P.cam = *cam;
    P.su = *su;
    if (ext) P.ext = *ext;
    P.object_id = 0;
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    if (ext && (ext->march_flags & RKFM_MARCH_TILE_CULL)) {
        const uint32_t tx = (out->width + 7u) / 8u, ty = (out->height + 7u) / 8u;
        if ((uint64_t)tx * ty > RKFM_NOALLOC_MAX_TILES)
            return rkf_set_error(RKFM_ERR_OUT_OF_RANGE, "march_noalloc: %ux%u frame has %llu tiles (max %u)",
                                 out->width, out->height, (unsigned long long)tx * ty, RKFM_NOALLOC_MAX_TILES);
        if (s->info.n_blocks > RKFM_NOALLOC_MAX_BLOCKS)
            return rkf_set_error(RKFM_ERR_OUT_OF_RANGE, "march_noalloc: scene has %u L1 blocks (max %u)",
                                 s->info.n_blocks, RKFM_NOALLOC_MAX_BLOCKS);
        rc = rkf_tile_camera_check(cam);
        if (rc) return rc;
        unsigned char *b = (unsigned char *)base;
        uint32_t *big_count = (uint32_t *)b;
        uint2 *tiles = (uint2 *)(b + kCountBytes);
        rkf::TileBig *big = (rkf::TileBig *)(b + kBigOffset);
        rkf::SparseDev D = s->dev();
        // tiles_clear_kernel zeroes the counter and every interval: nothing
        // of the previous call's contents is read
        rc = rkf_hip_check(rkf_launch_sparse_tiles(&D, cam, su->max_distance, out->width, out->height, tiles, nullptr, big,
                                                   big_count, true, st),
                           "march_noalloc tile intervals");
        if (rc) return rc;
        P.tiles = tiles;
        P.tiles_x = tx;
    }
    if (variant == 1 && !c->tuned_waves) {
        rc = rkf_hip_check(rkf_occupancy_waves(0, &c->tuned_waves), "occupancy");
        if (rc) return rc;
    }
    return rkf_hip_check(rkf_launch_sparse_march(&P, variant, c->tuned_waves, st), "march_noalloc launch");
}

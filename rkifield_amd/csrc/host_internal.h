// host_internal.h — context and buffer types private to librkf_march.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/rkf_march.h"
#include "kernels.h"

int rkf_set_error(int code, const char *fmt, ...);
int rkf_hip_check(hipError_t e, const char *what);
int rkf_fill_gbuf(const rkfm_gbuffer *out, rkf::GbufPtrs &g);

// Device buffer that grows by reallocation (GpuScene's buffers grow the same
// way, gpu_scene.rs:141-195).
struct DevBuf {
    void *ptr = nullptr;
    uint64_t cap = 0;
    uint64_t size = 0;
    int reserve(uint64_t want, bool keep);
    void release();
};

struct rkfm_sparse_scene {
    rkfm_ctx *owner = nullptr;
    rkfm_sparse_info info{};
    DevBuf l2, l1, dist, mat, aux, free_idx, free_bits, oct, cellw, roww;
    DevBuf block_cell;    // per L1 block: its flat L2 cell
    DevBuf edit_touched;  // one bit per slot, all zero between edit calls
    uint64_t n_free_rows = 0;
    rkf::SparseDev dev() const;
};

struct rkfm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint64_t memory_budget = 0;
    // GpuScene (reference layout)
    DevBuf pool, maps, objects, bvh, tile_idx, tile_counts, coarse;
    DevBuf ref_order;          // the tiled march's launch order (ref_order_n tiles)
    uint64_t ref_order_n = 0;
    uint64_t n_pool_voxels = 0, n_maps = 0;
    uint32_t n_objects = 0, n_bvh = 0, tiles_x = 0, tiles_y = 0;
    rkf_camera_uniforms cam{};
    rkf_scene_uniforms su{};
    rkfm_scene_ext ext{};
    rkf_coarse_field_uniforms coarse_u{};
    bool cam_set = false, scene_set = false, coarse_set = false;
    // Resident sparse scenes.
    std::vector<rkfm_sparse_scene *> scenes;
    // Persistent-wave count of the tuned sparse kernel (occupancy x CUs).
    uint32_t tuned_waves = 0;
    // Split sparse march: compact hit list and its counter.
    DevBuf march_hits, march_count;
    // Tile intervals of the last culled march: per-tile pairs, the large
    // rectangle records and their counter.
    DevBuf tile_iv, tile_big, tile_big_count, tile_cnt, tile_order;
    // Leading tile_iv entries (with tile_big_count) and tile_cnt entries (with
    // the 512 sort words before them) that the last plain march put back to
    // the cleared state, on stream clean_stream: the next culled view on that
    // stream bins into them without a clearing launch.
    uint64_t iv_clean = 0, cnt_clean = 0;
    hipStream_t clean_stream = nullptr;
    // Edits: op records (and their pinned staging, reusable once edit_ev
    // has fired), the listed bricks, their modified flags, the list count.
    DevBuf edit_ops, edit_list, edit_mod, edit_count;
    void *edit_host = nullptr;
    uint64_t edit_host_cap = 0;
    hipEvent_t edit_ev = nullptr;
    bool edit_staged = false;
    // Batched launch: device descriptors (rewritten per call, stream-ordered),
    // their host staging copies, the persistent work counter, the tuning.
    DevBuf batch_items, batch_begins, batch_counter;
    DevBuf batch_tiles;  // tile intervals of the culled views (layout: batch.hip)
    std::vector<unsigned char> batch_host_items;
    std::vector<uint32_t> batch_host_begins;
    rkfm_batch_tuning batch_tuning{};
    bool batch_tuned = false;
};

void rkf_sparse_release_all(rkfm_ctx *c);

// Batched launch internals (batch.hip), shared with the request server.
bool rkf_batch_valid_shape(const rkfm_batch_tuning &t);
rkfm_batch_tuning rkf_batch_default_tuning(const rkfm_ctx *c);
// Device descriptors and unit begins of a batch set; *total = work units.
// cull_views receives the views that ask for RKFM_MARCH_TILE_CULL; without
// it (the request server) any march flag is an error.
int rkf_batch_prepare(rkfm_ctx *c, const rkfm_batch_item *items, uint32_t n, const rkfm_batch_tuning &t,
                      std::vector<rkf::BatchDev> &dev, std::vector<uint32_t> &begins, uint32_t &total,
                      std::vector<uint32_t> *cull_views = nullptr);
// The camera checks of the tile-interval kernels (jitter, invertible basis).
extern "C" int rkf_tile_camera_check(const rkf_camera_uniforms *cam);
// Persistent-grid size: per_cu wavefronts on every CU of the current device.
int rkf_device_waves(uint32_t per_cu, uint32_t &waves);
// Looks up this device's copy of the allocation-free march's scratch array
// (noalloc.hip); called by rkfm_create with the device current.
int rkf_noalloc_init(int device);
// Enqueues the batch kernel (and the counter reset of the persistent
// schedule) over device copies of the descriptors and begins.
int rkf_batch_enqueue(const rkfm_batch_tuning &t, const rkf::BatchDev *items, const uint32_t *begins, uint32_t n,
                      uint32_t total, uint32_t *counter, uint32_t waves, hipStream_t st);

// 64-bit finaliser (splitmix64) used for the per-voxel material hash of the
// synthetic scenes; the same function runs on the device in sparse.hip.
static inline uint64_t rkf_mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// U{0..15} per (seed, object, global voxel coordinate).
static inline uint32_t rkf_voxel_material(uint64_t seed, uint32_t object, uint32_t gx, uint32_t gy, uint32_t gz) {
    uint64_t key = (uint64_t)gx | ((uint64_t)gy << 21) | ((uint64_t)gz << 42);
    return (uint32_t)(rkf_mix64(seed ^ rkf_mix64(key ^ ((uint64_t)object * 0xD6E8FEB86659FD93ull))) & 15u);
}
// Secondary material (6 bits) | blend weight (8 bits) << 8 of a voxel: a
// second hash of the same key with another salt.
static inline uint32_t rkf_voxel_blend(uint64_t seed, uint32_t object, uint32_t gx, uint32_t gy, uint32_t gz) {
    uint64_t key = (uint64_t)gx | ((uint64_t)gy << 21) | ((uint64_t)gz << 42);
    uint64_t h = rkf_mix64(seed ^ rkf_mix64(key ^ ((uint64_t)object * 0xD6E8FEB86659FD93ull)) ^ 0xA24BAED4963EE407ull);
    return (uint32_t)((h >> 8) & 0x3Fu) | ((uint32_t)((h >> 16) & 0xFFu) << 8);
}

// kernels.h — launch parameter records shared by the host code and kernels.
#pragma once

#include "device_common.h"

namespace rkf {

// Drop-in march over the reference layouts (march_ref.hip).
struct RefParams {
    const rkf_voxel_sample *pool;
    uint64_t n_pool_voxels;
    const uint32_t *maps;
    uint64_t n_maps;
    const rkf_gpu_object *objects;
    const rkf_bvh_node *bvh;
    uint32_t n_bvh;
    const uint32_t *tile_idx;
    const uint32_t *tile_counts;
    uint32_t n_tiles;
    const uint32_t *tile_order;  // tiled mode: tiles by descending object count (nullptr: row-major launch)
    const uint16_t *coarse;  // nullptr: no coarse field
    rkf_coarse_field_uniforms cu;
    rkf_camera_uniforms cam;
    rkf_scene_uniforms su;
    uint32_t eps_mode;
    uint32_t mode;  // 0 tiled, 1 bvh, 2 brute
    GbufPtrs g;
};

// Producers (producers.hip).
struct TileCullParams {
    const rkf_gpu_object *objects;
    uint32_t num_objects;
    rkf_camera_uniforms cam;
    uint32_t width, height, tiles_x, tiles_y;
    uint32_t *idx;
    uint32_t *counts;
};
struct CoarseParams {
    const rkf_gpu_object *objects;
    uint32_t n_objects;
    float origin[3];
    uint32_t dims[3];
    float voxel_size;
    uint16_t *cells;
};

// Sparse two-level scene as the kernels see it (sparse.hip).
constexpr uint32_t kFreeCell = 0xFFFFFFFFu;
constexpr uint32_t kOctCap = 15u;          // octant extents saturate at 15 cells
constexpr uint32_t kNoSlot = 0xFFFFFFFDu;  // "slot not prefetched"
struct SparseDev {
    float origin[3];
    float vs;
    float inv_vs;     // 1 / vs rounded to float: g = (p - origin) * inv_vs - 0.5
    uint32_t cells[3];
    uint32_t nb[3];   // bricks per axis = cells * 8
    int32_t tot[3];   // voxels per axis = cells * 64
    const uint32_t *l2;        // per L2 cell: L1 block or EMPTY
    const uint32_t *l1;        // per block: 512 slots (brick x + 8y + 64z in cell)
    const uint16_t *dist;      // per slot: 512 binary16 distances
    const uint8_t *mat;        // per slot: 512 material ids (6 bits)
    const uint16_t *aux;       // per slot: 512 (secondary | blend << 8), or null
    const uint32_t *free_idx;  // per L2 cell: kFreeCell or row of free_bits
    const uint64_t *free_bits; // 8 words per row: bit set = brick NOT free
    const uint32_t *oct;       // per L2 cell: 8 octant extents, 4 bits each
    const uint64_t *roww;      // per row brick: octant extents (4 bits each) << 32
                               // | L1 slot (EMPTY in rows past n_blocks)
    const uint64_t *cellw;     // per L2 cell, what one DDA step needs in one
                               // load: free row << 32 | octant extents
    uint64_t n_cells;
    uint32_t n_blocks;         // rows below n_blocks are the L1 blocks
    const uint32_t *block_cell;  // per L1 block: its flat L2 cell
};

struct SparseMarchParams {
    SparseDev S;
    rkf_camera_uniforms cam;
    rkf_scene_uniforms su;
    rkfm_scene_ext ext;
    GbufPtrs g;
    uint32_t object_id;
    uint32_t tiles_x;          // 8x8 tiles per row of `tiles`
    const uint2 *tiles;        // per tile (t_lo, t_hi) float bits, or null
    const uint32_t *order;     // plain kernel: tiles in launch order, or null (row-major)
    // Plain kernel, optional: after marching its tile, each workgroup puts
    // the tile's interval (and count) back to the cleared state, and the
    // first workgroup clears the sort bins and the large-rectangle count, so
    // the next culled view's binning needs no clearing launch. Null: left.
    uint2 *reset_iv;
    uint32_t *reset_cnt, *reset_bins, *reset_big;
};

// A cell whose tile rectangle is too large for one lane (tile binning).
struct TileBig {
    uint32_t x0, x1, y0, y1;
    float tn, tf;
};
constexpr uint32_t kTileInlineMax = 16u;  // tiles one lane bins itself

// One batch of a batched launch (batch.hip): the batch's scene and
// parameters plus its range of 64-lane work units.
struct BatchDev {
    SparseMarchParams P;  // scene, camera, uniforms, extension, outputs
    uint32_t kind, n, unit_begin, n_units;
    uint32_t tile_w, tile_h, tiles_x, pad;
    const float *in;
    float *vis;
    rkfm_point_result *res;
    const rkfm_sphere_cast *casts;
    rkfm_cast_result *cres;
};

}  // namespace rkf

extern "C" {
hipError_t rkf_launch_ref_march(const rkf::RefParams *P, hipStream_t stream);
hipError_t rkf_launch_copy(void *dst, const void *src, uint64_t bytes, hipStream_t stream);
hipError_t rkf_launch_gather16(void *dst, const void *src, uint64_t n_lines, uint64_t n, hipStream_t stream);
hipError_t rkf_launch_tile_cull(const rkf::TileCullParams *P, hipStream_t stream);
hipError_t rkf_launch_ref_tile_order(const uint32_t *counts, uint32_t n_tiles, uint32_t *order, hipStream_t stream);
hipError_t rkf_launch_coarse(const rkf::CoarseParams *P, hipStream_t stream);
hipError_t rkf_launch_sparse_march(const rkf::SparseMarchParams *P, uint32_t variant, uint32_t waves,
                                   hipStream_t stream);
hipError_t rkf_occupancy_waves(uint32_t which, uint32_t *waves);
uint64_t rkf_hit_record_bytes();
hipError_t rkf_launch_sparse_contacts(const rkf::SparseDev *S, const rkfm_body *bodies, uint32_t n, rkfm_contact *out,
                                      uint32_t *counts, hipStream_t stream);
hipError_t rkf_launch_sparse_march_split(const rkf::SparseMarchParams *P, void *hits, uint32_t *count,
                                         uint32_t shade_waves, hipStream_t stream);
hipError_t rkf_launch_sparse_query(const rkf::SparseDev *S, const float *xyz, uint32_t n, rkfm_point_result *out,
                                   hipStream_t stream);
hipError_t rkf_launch_sparse_cast(const rkf::SparseDev *S, const rkfm_sphere_cast *casts, uint32_t n,
                                  rkfm_cast_result *out, hipStream_t stream);
// Per-tile primary-ray intervals: clear, bin every L1 block's cell, then the
// cells whose rectangle exceeds kTileInlineMax tiles (big holds up to
// n_blocks records, big_count one counter). cnt (or null): per tile, the
// number of cells binned into it; the 512 words before cnt[0] must exist
// (rkf_launch_sparse_tile_order's bins, zeroed here).
hipError_t rkf_launch_sparse_tiles(const rkf::SparseDev *S, const rkf_camera_uniforms *cam, float max_distance,
                                   uint32_t width, uint32_t height, uint2 *tiles, uint32_t *cnt, rkf::TileBig *big,
                                   uint32_t *big_count, bool clear, hipStream_t stream);
// Tiles by descending cell count (order holds n_tiles entries; cnt as filled
// by rkf_launch_sparse_tiles, with its bins before it).
hipError_t rkf_launch_sparse_tile_order(uint32_t *cnt, uint32_t n_tiles, uint32_t *order, hipStream_t stream);
}

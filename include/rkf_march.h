/*
 * rkf_march.h — C ABI of the MI355X (gfx950) SDF ray-march library.
 *
 * Drop-in replacement for rkf-render's ray-march pass
 * (crates/rkf-render/src/ray_march.rs:29-118, kernel
 * crates/rkf-render/shaders/ray_march.wgsl). The byte layouts of every input
 * record are the reference's own (static-asserted below and in the tests):
 *
 *   rkf_voxel_sample       crates/rkf-core/src/voxel.rs:15-20          8 B
 *   rkf_gpu_object         crates/rkf-render/src/gpu_object.rs:41-95   256 B
 *   rkf_camera_uniforms    crates/rkf-render/src/camera.rs:194-213     144 B
 *   rkf_scene_uniforms     crates/rkf-render/src/gpu_scene.rs:21-43    16 B
 *   rkf_coarse_field_uniforms crates/rkf-render/src/coarse_field.rs:31-42 48 B
 *   rkf_bvh_node           crates/rkf-render/src/gpu_bvh.rs:24-41      32 B
 *
 * Conventions
 *   - Every function returns an int status: RKFM_OK (0) or a negative error
 *     code; rkfm_last_error_string() gives the message of the last failure
 *     on the calling thread.
 *   - A context (rkfm_ctx) owns its device memory and is thread-affine: use
 *     one host thread per context.
 *   - Host pointers passed to upload/set functions are copied before return.
 *   - G-buffer and query output pointers are device pointers owned by the
 *     caller; launches are asynchronous on the given stream (NULL = the
 *     context's own stream).
 */
#ifndef RKF_MARCH_H
#define RKF_MARCH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Status codes                                                              */
/* ------------------------------------------------------------------------ */
#define RKFM_OK 0
#define RKFM_ERR_INVALID_ARG (-1)  /* null pointer, zero size, bad enum      */
#define RKFM_ERR_OUT_OF_RANGE (-2) /* slot/offset/index beyond an upload     */
#define RKFM_ERR_OOM (-3)          /* device allocation failed / over budget */
#define RKFM_ERR_DEVICE (-4)       /* HIP runtime error                      */
#define RKFM_ERR_LAYOUT (-5)       /* inconsistent layout (e.g. map dims)    */
#define RKFM_ERR_NOT_READY (-6)    /* required input never uploaded          */
#define RKFM_ERR_BUSY (-7)         /* request queue full                     */
#define RKFM_ERR_FORMAT (-8)       /* malformed or unsupported asset data    */

/* ------------------------------------------------------------------------ */
/* Reference constants                                                       */
/* ------------------------------------------------------------------------ */
#define RKF_EMPTY_SLOT 0xFFFFFFFFu    /* brick_map.rs:31                      */
#define RKF_INTERIOR_SLOT 0xFFFFFFFEu /* brick_map.rs:37                      */
#define RKF_BVH_INVALID 0xFFFFFFFFu   /* ray_march.wgsl:27                    */
#define RKF_BRICK_DIM 8u              /* constants.rs:10                      */
#define RKF_TILE_SIZE 16u             /* ray_march.wgsl:47                    */
#define RKF_TILE_MAX_OBJECTS 32u      /* ray_march.wgsl:48                    */
#define RKF_SDF_TYPE_NONE 0u
#define RKF_SDF_TYPE_ANALYTICAL 1u
#define RKF_SDF_TYPE_VOXELIZED 2u

/* ------------------------------------------------------------------------ */
/* Reference layouts                                                         */
/* ------------------------------------------------------------------------ */

/* voxel.rs:15-20. word0 = f16 distance (bits 0-15) | material (16-21) |
 * secondary material (22-27); word1 byte 3 = blend weight. */
typedef struct rkf_voxel_sample {
    uint32_t word0;
    uint32_t word1;
} rkf_voxel_sample;

/* gpu_object.rs:41-95 (WGSL view ray_march.wgsl:60-97). inverse_world is
 * column-major: inverse_world[c*4 + r]. */
typedef struct rkf_gpu_object {
    float inverse_world[16];
    float aabb_min[4];
    float aabb_max[4];
    uint32_t brick_map_offset;
    uint32_t brick_map_dims[3];
    float voxel_size;
    uint32_t material_id;
    uint32_t sdf_type;
    uint32_t blend_mode;
    float blend_radius;
    float sdf_params[4];
    float accumulated_scale[3];
    uint32_t lod_level;
    uint32_t object_id;
    uint32_t primitive_type;
    float geometry_aabb_min[3];
    float geometry_aabb_max[3];
    float _padding[15];
} rkf_gpu_object;

/* camera.rs:194-213. prev_vp column-major. */
typedef struct rkf_camera_uniforms {
    float position[4];
    float forward[4];
    float right[4];
    float up[4];
    float resolution[2];
    float jitter[2];
    float prev_vp[16];
} rkf_camera_uniforms;

/* gpu_scene.rs:21-43; defaults 0/128/100.0/0.001 (gpu_scene.rs:34-43). */
typedef struct rkf_scene_uniforms {
    uint32_t num_objects;
    uint32_t max_steps;
    float max_distance;
    float hit_threshold;
} rkf_scene_uniforms;

/* coarse_field.rs:31-42. */
typedef struct rkf_coarse_field_uniforms {
    float origin_cam_rel[4];
    uint32_t dims[4];
    float voxel_size;
    float inv_voxel_size;
    float _pad[2];
} rkf_coarse_field_uniforms;

/* gpu_bvh.rs:24-41 (WGSL ray_march.wgsl:99-108). left == RKF_BVH_INVALID
 * marks a leaf whose right_or_object is the object index. */
typedef struct rkf_bvh_node {
    float aabb_min[3];
    uint32_t left;
    float aabb_max[3];
    uint32_t right_or_object;
} rkf_bvh_node;

/* ------------------------------------------------------------------------ */
/* Extension records (no reference counterpart; zero-initialised = reference */
/* behaviour)                                                                */
/* ------------------------------------------------------------------------ */

#define RKFM_EPS_ABSOLUTE 0u /* hit if d < hit_threshold (ray_march.wgsl:966) */
#define RKFM_EPS_RELATIVE 1u /* hit if d < hit_threshold * t                  */

/* Strategy override for the drop-in kernel: AUTO follows ray_march.wgsl:1056-1063
 * (tiled if tile lists are set, else BVH if nodes are set, else brute). */
#define RKFM_STRATEGY_AUTO 0u
#define RKFM_STRATEGY_TILED 1u
#define RKFM_STRATEGY_BVH 2u
#define RKFM_STRATEGY_BRUTE 3u

/* Sparse scenes: bound every 8x8-pixel tile's primary rays by the occupied
 * L2 cells its frustum can meet (rkfm_sparse_tile_intervals). A ray starts
 * at its tile's t_lo (one skip, prev_t = t_lo, prev_dist = 8 vs, like a
 * free-space skip) and ends as a miss past t_hi; a tile that meets no
 * occupied cell misses after that one skip. Outside [t_lo, t_hi] the field
 * along the ray is the free value, so only the path through free space
 * changes, as with the free-space DDA. */
#define RKFM_MARCH_TILE_CULL 1u

typedef struct rkfm_scene_ext {
    uint32_t hit_eps_mode;   /* RKFM_EPS_*                                  */
    uint32_t strategy;       /* RKFM_STRATEGY_*                             */
    uint32_t shadow_enable;  /* 1: trace one soft-shadow ray per hit        */
    uint32_t march_flags;    /* RKFM_MARCH_* (sparse scenes)                */
    float light_dir[3];      /* unit vector toward the light                */
    float shadow_max_dist;   /* 0 => 25.0  (shade_common.wgsl:276)          */
    float shadow_k;          /* 0 => 16.0  (shade_common.wgsl:275)          */
    float _reserved1[3];
} rkfm_scene_ext;

/* Per-pixel debug record (optional plane). All integer fields are exact. */
typedef struct rkfm_debug_px {
    float normal[3];     /* f32 normal before the f16 G-buffer store        */
    uint32_t brick_slot; /* pool slot of the voxel holding the hit point    */
                         /* (RKF_EMPTY_SLOT on miss / sentinel)             */
    uint32_t voxel_index;/* x + 8y + 64z inside that brick (0xFFFFFFFF miss) */
    uint32_t cell_index; /* flat L2 cell (sparse scenes) or object index    */
    uint32_t steps;      /* march loop iterations (incl. DDA skips)         */
    uint32_t evals;      /* SDF evaluations made (incl. bisection)          */
    uint32_t shadow_steps; /* shadow-ray loop iterations (0 if none)        */
    uint32_t shadow_evals; /* shadow-ray SDF evaluations                    */
    uint32_t _pad[2];
} rkfm_debug_px;

/* G-buffer planes, row-major (x fastest), in the reference formats
 * (gbuffer.rs:52-58): position rgba32f, normal rgba16f, material r32u,
 * motion rgba32f. Optional planes may be NULL. All are device pointers. */
typedef struct rkfm_gbuffer {
    void *position;     /* float[4] per pixel                               */
    void *normal;       /* uint16_t[4] (binary16) per pixel                 */
    void *material;     /* uint32_t per pixel                               */
    void *motion;       /* float[4] per pixel                               */
    void *debug;        /* rkfm_debug_px per pixel, optional                */
    void *shadow;       /* float per pixel (visibility), optional           */
    uint32_t width;
    uint32_t height;
} rkfm_gbuffer;

/* ------------------------------------------------------------------------ */
/* Context                                                                    */
/* ------------------------------------------------------------------------ */
typedef struct rkfm_ctx rkfm_ctx;

typedef struct rkfm_config {
    int device;               /* HIP device ordinal                          */
    uint32_t flags;           /* reserved, 0                                 */
    uint64_t memory_budget;   /* bytes the served loop may hold; 0 = no cap  */
} rkfm_config;

const char *rkfm_last_error_string(void);
const char *rkfm_version(void);

/* RayMarchPass::new (ray_march.rs:42-80): creates device state and loads the
 * kernels; no pipeline compilation happens at run time. */
int rkfm_create(const rkfm_config *cfg, rkfm_ctx **out);
int rkfm_destroy(rkfm_ctx *ctx);
int rkfm_sync(rkfm_ctx *ctx);
/* The context's stream (hipStream_t) for callers that want to chain work. */
void *rkfm_stream(rkfm_ctx *ctx);

/* GpuScene uploads (render.rs:106-137, gpu_scene.rs:141-195). The device
 * buffers grow as needed; first_slot / offset address the destination.
 * Voxels and map entries are indexed in 32 bits, as the reference's u32
 * indices are: RKFM_ERR_OUT_OF_RANGE past slot 2^23 - 2 (a 32 GiB pool) or
 * map entry 2^32 - 2. */
int rkfm_upload_brick_pool(rkfm_ctx *ctx, const rkf_voxel_sample *bricks,
                           uint64_t n_bricks, uint64_t first_slot);
int rkfm_upload_brick_maps(rkfm_ctx *ctx, const uint32_t *entries,
                           uint64_t n, uint64_t offset);
int rkfm_upload_objects(rkfm_ctx *ctx, const rkf_gpu_object *objects,
                        uint32_t n);
int rkfm_upload_bvh(rkfm_ctx *ctx, const rkf_bvh_node *nodes, uint32_t n);
int rkfm_set_camera(rkfm_ctx *ctx, const rkf_camera_uniforms *camera);
int rkfm_set_scene(rkfm_ctx *ctx, const rkf_scene_uniforms *scene,
                   const rkfm_scene_ext *ext /* NULL = reference */);
/* Per-tile object lists as produced by tile_object_cull.wgsl:117-204:
 * idx holds tiles_x*tiles_y*32 u32, counts tiles_x*tiles_y u32. Passing
 * tiles_x == 0 clears the lists (strategy AUTO then falls back to BVH/brute).
 * Here and in rkfm_build_tiles the context also orders the tiles for the
 * tiled march's launch (descending object count; no effect on results). */
int rkfm_set_tiles(rkfm_ctx *ctx, const uint32_t *idx, const uint32_t *counts,
                   uint32_t tiles_x, uint32_t tiles_y);
/* Coarse field (coarse_field.rs:240-307): dims.x*dims.y*dims.z binary16
 * cells, x fastest. NULL cells clears the field (every lookup returns 0). */
int rkfm_set_coarse_field(rkfm_ctx *ctx, const uint16_t *f16_cells,
                          const rkf_coarse_field_uniforms *uniforms);

/* Producers of the march inputs, on the GPU (tile_object_cull.wgsl:117-204 and
 * CoarseField::populate coarse_field.rs:240-266). Tile lists are ordered by
 * ascending object index, keeping the first 32. Both launch asynchronously on
 * the context's stream (rkfm_stream): an rkfm_dispatch on another stream must
 * be ordered after them (rkfm_sync, or an event recorded on that stream). */
int rkfm_build_tiles(rkfm_ctx *ctx, uint32_t width, uint32_t height);
int rkfm_build_coarse_field(rkfm_ctx *ctx, const float origin[3],
                            const uint32_t dims[3], float voxel_size);
/* Read back what the producers built (host pointers; sizes in elements). */
int rkfm_read_tiles(rkfm_ctx *ctx, uint32_t *idx, uint32_t *counts,
                    uint64_t n_tiles);
int rkfm_read_coarse_field(rkfm_ctx *ctx, uint16_t *cells, uint64_t n_cells,
                           rkf_coarse_field_uniforms *uniforms);

/* RayMarchPass::dispatch (ray_march.rs:96-118): one thread per pixel of
 * out->width x out->height, 8x8 pixel tiles per wavefront. */
int rkfm_dispatch(rkfm_ctx *ctx, const rkfm_gbuffer *out, void *hip_stream);

/* ------------------------------------------------------------------------ */
/* Host-side scene building (restates rkf-core voxelize_object.rs:53-182)    */
/* ------------------------------------------------------------------------ */

/* Analytic primitive record used by the scene builders. type: 0 sphere
 * (p0 = radius), 1 box (p0..p2 = half extents), 2 capsule (p0 = radius,
 * p1 = half height along local y), 3 fBm heightfield terrain (world
 * aligned, rot unused): d = p2 * ((y - centre.y) - h(x - centre.x,
 * z - centre.z)) with h = p0 * sum_{i<6} 0.5^i * vnoise(seed + i, f_i x,
 * f_i z), f_i = p1 * 2^i, vnoise the smoothstep-interpolated value noise of
 * a splitmix64 lattice hash in [-1, 1) (the same height function terrain.rs
 * evaluates as y - h, with fBm detail instead of sine layers). rot is a
 * row-major 3x3 world->local rotation applied after subtracting centre.
 * group: primitives with the same group are blended with smin(k)
 * (sdf.rs:48-51), groups are unioned by min. */
typedef struct rkfm_prim {
    float centre[3];
    uint32_t type;
    float params[4];
    float rot[9];
    uint32_t group;
    uint32_t material; /* used when the scene has no per-voxel material hash */
    uint32_t seed;     /* noise seed (type 3)                                */
} rkfm_prim;

/* voxelize_sdf (voxelize_object.rs:53-157) of one sphere |p| - r with the
 * AABB +-(r + margin): writes the brick map (dims product entries, slots
 * numbered first_slot, first_slot+1, ... in z-y-x order as Pool::allocate_range
 * yields them) and the bricks (4 KiB each). Per-voxel material is
 * hash(seed, object, voxel) & 15 when mat_seed != 0, else `material`.
 * Call with bricks == NULL to size: *n_bricks and dims are filled. */
int rkfm_voxelize_sphere(float radius, float margin, float voxel_size,
                         uint32_t object_index, uint64_t mat_seed,
                         uint32_t material, uint32_t first_slot,
                         uint32_t dims[3], uint32_t *map_entries,
                         rkf_voxel_sample *bricks, uint64_t *n_bricks);

/* Brick classification and voxel contents for rkfm_voxelize_prim.
 *   RKFM_VOX_GEOMETRY_FIRST: voxelize_to_geometry's classification
 *     (voxelize_object.rs:217-385): a brick whose centre distance is below
 *     -1.8 brick extents is INTERIOR_SLOT; a narrow-band brick whose 512
 *     voxel-centre distances are all <= 0 is INTERIOR_SLOT, all > 0
 *     EMPTY_SLOT; the other narrow-band bricks are stored (VoxelSample words
 *     as voxelize_sdf writes them) and take consecutive slots in z-y-x order
 *     (the reference's pool frees the slots of the bricks it drops, leaving
 *     holes; here the numbering is compact).
 *   RKFM_VOX_BLEND (with GEOMETRY_FIRST): each stored voxel also carries a
 *     secondary material and blend weight (VoxelSample::new_blended,
 *     voxel.rs:43-55) from a second hash of its key: secondary =
 *     (h >> 8) & 63, blend = (h >> 16) & 255. */
#define RKFM_VOX_GEOMETRY_FIRST 1u
#define RKFM_VOX_BLEND 2u

/* voxelize_sdf (flags 0) or the classification above of one primitive (type
 * 0-2; centre and rot place it in the object's local space) over the AABB
 * centre +- (bounding radius + margin); per-voxel material and sizing call
 * as rkfm_voxelize_sphere. */
int rkfm_voxelize_prim(const rkfm_prim *prim, float margin, float voxel_size,
                       uint32_t flags, uint32_t object_index, uint64_t mat_seed,
                       uint32_t material, uint32_t first_slot, uint32_t dims[3],
                       uint32_t *map_entries, rkf_voxel_sample *bricks,
                       uint64_t *n_bricks);

/* The 5-sphere smooth union of crates/rkf-core/examples/gen_test_rkf.rs:22-127
 * (vs 0.03, k 0.15, material 6), voxelised the same way. */
int rkfm_voxelize_test_cross(uint32_t dims[3], uint32_t *map_entries,
                             rkf_voxel_sample *bricks, uint64_t *n_bricks);

/* Camera looking from `position` at `target` with the reference fly-camera
 * parameterisation (camera.rs:75-93, uniforms camera.rs:161-184). */
int rkfm_camera_look_at(const float position[3], const float target[3],
                        float fov_degrees, uint32_t width, uint32_t height,
                        rkf_camera_uniforms *out);

/* Median-split BVH over object AABBs (aabb_min/aabb_max of the objects).
 * nodes must hold 2*n-1 entries; *n_nodes receives the count. */
int rkfm_build_bvh(const rkf_gpu_object *objects, uint32_t n,
                   rkf_bvh_node *nodes, uint32_t *n_nodes);

/* ------------------------------------------------------------------------ */
/* Sparse world scenes: two-level index (L2 cells of 8^3 bricks -> L1 brick   */
/* slots) over a structure-of-arrays brick pool (f16 distance plane + u8     */
/* material plane), resident in HBM. No reference counterpart: it holds      */
/* scenes the reference's 1 GiB pool binding (context.rs:46) cannot.         */
/* ------------------------------------------------------------------------ */
typedef struct rkfm_sparse_scene rkfm_sparse_scene;

typedef struct rkfm_sparse_params {
    float voxel_size;       /* vs; brick extent = 8 vs, L2 cell = 64 vs      */
    float smin_k;           /* group blend radius (0 = hard min)           */
    uint64_t mat_seed;      /* per-voxel material hash seed (0 = prim mat) */
    float bounds_min[3];    /* world region to voxelise                    */
    float bounds_max[3];
    uint64_t max_bricks;    /* capacity cap (0 = no cap beyond memory)     */
} rkfm_sparse_params;

typedef struct rkfm_sparse_info {
    float origin[3];        /* world position of global voxel corner 0     */
    float voxel_size;
    uint32_t cells[3];      /* L2 grid dimensions                          */
    uint32_t n_blocks;      /* allocated L1 blocks (occupied L2 cells)     */
    uint64_t n_bricks;      /* allocated bricks                            */
    uint64_t device_bytes;  /* HBM held by the scene                       */
} rkfm_sparse_info;

/* Voxelise primitives on the GPU into a new resident sparse scene (at most
 * 32768 L2 cells per axis). */
int rkfm_sparse_build(rkfm_ctx *ctx, const rkfm_prim *prims, uint32_t n_prims,
                      const rkfm_sparse_params *params,
                      rkfm_sparse_scene **out);
/* Build a sparse scene from a reference dense brick map + AoS bricks (one
 * object, translation only: `origin` = world position of the map's grid
 * corner; at most 262144 bricks per axis). Used to stream .rkf assets and
 * fixtures into resident scenes. */
int rkfm_sparse_from_dense(rkfm_ctx *ctx, const uint32_t *map,
                           const uint32_t dims[3],
                           const rkf_voxel_sample *bricks, uint64_t n_bricks,
                           float voxel_size, const float origin[3],
                           rkfm_sparse_scene **out);
/* Reference scene -> resident sparse scene (AoS -> SoA; the renderer's
 * per-frame pool and object upload, render.rs:106-137, becomes a one-time
 * conversion). Resamples the reference-layout scene uploaded to this context
 * (brick pool, brick maps, objects and su.num_objects: what rkfm_dispatch
 * marches) onto a sparse scene of voxel size vs whose grid corner is
 * bounds_min. The field at a point is the minimum of evaluate_object
 * (ray_march.wgsl:600-634: inverse_world, analytic primitive or trilinear
 * voxel field with its sentinels, times min(accumulated_scale)) over the
 * objects whose AABB widened by 3 bricks meets the point's L2 cell, in
 * ascending index (the AABB must bound where an object's distance falls
 * below 3 bricks, as the coarse field and tile cull assume). A brick is
 * stored iff |field at its centre| < 1.8 brick extents (voxelize_object.rs:
 * 86,104); each voxel stores binary16(field at its centre), the material
 * evaluate_object returns for the first object reaching the minimum, and (in
 * an aux plane, present when some voxel has them) that object's nearest
 * voxel's secondary material and blend weight. */
int rkfm_sparse_from_objects(rkfm_ctx *ctx, float voxel_size,
                             const float bounds_min[3], const float bounds_max[3],
                             rkfm_sparse_scene **out);
/* The scene's aux plane (n_bricks * 512 u16: secondary | blend << 8) when it
 * has one (*present = 1); aux may be NULL to ask. */
int rkfm_sparse_download_aux(rkfm_ctx *ctx, const rkfm_sparse_scene *scene,
                             uint16_t *aux, uint32_t *present);
int rkfm_sparse_info_get(const rkfm_sparse_scene *scene,
                         rkfm_sparse_info *info);
int rkfm_sparse_destroy(rkfm_ctx *ctx, rkfm_sparse_scene *scene);
/* Host copies of the index and pool (sizes from rkfm_sparse_info): l2 has
 * cells.x*cells.y*cells.z u32, l1 n_blocks*512 u32, dist n_bricks*512
 * binary16, mat n_bricks*512 u8. Any pointer may be NULL. */
int rkfm_sparse_download(rkfm_ctx *ctx, const rkfm_sparse_scene *scene,
                         uint32_t *l2, uint32_t *l1, uint16_t *dist,
                         uint8_t *mat);

/* March a sparse scene: ray generation, sphere tracing with the reference
 * step rules, multi-level DDA over free space, normals, G-buffer write and
 * optional shadow rays. variant 0 = one lane per pixel, 8x8 pixels per
 * wavefront, hardware-scheduled; variant 1 = persistent wavefronts that
 * refill finished lanes; variant 2 = primary pass + compacted hit list +
 * shading/shadow pass (the context keeps a 36 B-per-pixel hit list). All
 * variants give the same results, bit for bit. With RKFM_MARCH_TILE_CULL the
 * context keeps the per-tile intervals (8 B per 8x8 tile) and, for variant 0
 * on frames of at most 65,536 tiles, a launch order of the tiles (most
 * occupied cells first, 8 B per tile + 2 KiB); the order changes only which
 * workgroup marches which tile, never a result. */
int rkfm_sparse_march(rkfm_ctx *ctx, const rkfm_sparse_scene *scene,
                      const rkf_camera_uniforms *camera,
                      const rkf_scene_uniforms *su, const rkfm_scene_ext *ext,
                      const rkfm_gbuffer *out, uint32_t variant,
                      void *hip_stream);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc). rkfm_sparse_march (variants 0 and 1 only) with the scratch of
 * RKFM_MARCH_TILE_CULL (per-tile ray intervals, large-rectangle list and its
 * counter) in a fixed device array of RKFM_NOALLOC_SCRATCH_BYTES that the
 * library holds on each device until the process ends, instead of context
 * buffers that grow: it calls no allocating routine. Culled frames are
 * limited to RKFM_NOALLOC_MAX_TILES 8x8 tiles and scenes to
 * RKFM_NOALLOC_MAX_BLOCKS L1 blocks (RKFM_ERR_OUT_OF_RANGE beyond). Every
 * call on one device shares that device's array: calls that overlap on one
 * device (other streams, contexts or threads) race. The results equal
 * rkfm_sparse_march's bit for bit. */
#define RKFM_NOALLOC_SCRATCH_BYTES 8389120u /* 256 + 262144 * 8 + 262144 * 24 */
#define RKFM_NOALLOC_MAX_TILES 262144u      /* e.g. 4096 x 4096 pixels       */
#define RKFM_NOALLOC_MAX_BLOCKS 262144u
int rkfm_sparse_march_noalloc(rkfm_ctx *ctx, const rkfm_sparse_scene *scene,
                              const rkf_camera_uniforms *camera,
                              const rkf_scene_uniforms *su,
                              const rkfm_scene_ext *ext, const rkfm_gbuffer *out,
                              uint32_t variant, void *hip_stream);

/* Per-tile ray intervals used by RKFM_MARCH_TILE_CULL: tiles_dev receives
 * ceil(width/8) * ceil(height/8) pairs (t_lo, t_hi) of floats, row-major.
 * For each L2 cell holding an L1 block, its box widened by 9 vs (by 25 vs on
 * the sides that lie on the grid boundary, where samples clamp) is projected
 * with the camera basis; every tile whose pixel rays can meet the box takes
 * t_lo = min(distance from the camera to the box) and t_hi = max(distance
 * to the box's farthest corner) over those cells. Boxes beyond
 * su->max_distance are skipped; a tile that meets none keeps
 * (+inf, 0). Requires |jitter| <= 0.5. */
int rkfm_sparse_tile_intervals(rkfm_ctx *ctx, const rkfm_sparse_scene *scene,
                               const rkf_camera_uniforms *camera,
                               const rkf_scene_uniforms *su, uint32_t width,
                               uint32_t height, float *tiles_dev,
                               void *hip_stream);

/* No counterpart in the reference: the launch order rkfm_sparse_march uses
 * for a culled view (variant 0, at most 65,536 tiles). counts_dev receives,
 * per tile (row-major as above), the number of L2 cells binned into it;
 * order_dev a permutation of the tile indices by descending
 * min(count, 255) (within one count, any order). Frames above 65,536 tiles:
 * RKFM_ERR_OUT_OF_RANGE. Requires |jitter| <= 0.5. */
int rkfm_sparse_tile_order(rkfm_ctx *ctx, const rkfm_sparse_scene *scene,
                           const rkf_camera_uniforms *camera,
                           const rkf_scene_uniforms *su, uint32_t width,
                           uint32_t height, uint32_t *counts_dev,
                           uint32_t *order_dev, void *hip_stream);

/* Point queries (SdfQueryable semantics, sdf_collision.rs:38-56; result
 * fields as SpatialQueryResult automation/types.rs:151-158 plus gradient). */
typedef struct rkfm_point_result {
    float distance;
    uint32_t material;
    uint32_t inside;
    float gradient[3];
} rkfm_point_result;

int rkfm_sparse_query_points(rkfm_ctx *ctx, const rkfm_sparse_scene *scene,
                             const float *xyz_dev, uint32_t n,
                             rkfm_point_result *out_dev, void *hip_stream);

/* Sphere casts (no reference counterpart: the reference answers point
 * queries only, sdf_collision.rs:38-56): a ball of radius r swept from
 * origin along the unit vector dir is sphere-traced against the stored field
 * with a conservative distance D(p) at its centre p:
 *   L(p) = a lower bound on the distance from p to any stored (non-EMPTY)
 *          brick, from the free-space structures the DDA uses: p outside
 *          the grid -> the distance to the grid box; p in a free L2 cell
 *          whose 8 octant extents have minimum E -> the distance to the faces
 *          of the free box of cells [q - E + 1, q + E) around it; p in a free
 *          brick (minimum brick extent E_b) -> likewise over bricks; else 0.
 *          The surface lies in stored bricks only (narrow band), so the true
 *          distance is >= L(p); where the stored field is capped (EMPTY reads
 *          8 vs) L lets a ball cross free space in large steps. D never
 *          exceeds the distance to the surface, so a contact is never late;
 *          for r above 8 vs it can come up to the narrow band (1.8 bricks)
 *          early, where L stops at the stored bricks around the surface.
 *   D(p) = L(p) when L(p) >= 16 vs (the sample there reads only EMPTY
 *          corners), else max(stepping sample at p, L(p)).
 * Starting at t = 0, a cast hits at the first t with D(p) - r < 1e-3; else
 * t += max(D(p) - r, 5e-4), at most 1024 iterations, a miss past max_t.
 * A hit's normal is the 6-tap central difference (2 vs, gradient sentinels)
 * at the contact centre p = origin + t dir, normalised ((0, 1, 0) when
 * shorter than 1e-10); the contact point is p - r * normal. */
typedef struct rkfm_sphere_cast {
    float origin[3];
    float radius;
    float dir[3];        /* unit vector                                    */
    float max_t;
} rkfm_sphere_cast;

typedef struct rkfm_cast_result {
    float t;             /* centre parameter of the contact (miss: max_t)  */
    uint32_t hit;
    uint32_t steps;      /* loop iterations                                */
    float normal[3];     /* at the contact centre (miss: 0)                */
    float distance;      /* D at the last centre evaluated                  */
    uint32_t _pad;
} rkfm_cast_result;

/* casts and out are device pointers (n records each). */
int rkfm_sparse_sphere_cast(rkfm_ctx *ctx, const rkfm_sparse_scene *scene,
                            const rkfm_sphere_cast *casts_dev, uint32_t n,
                            rkfm_cast_result *out_dev, void *hip_stream);
/* Contact generation (rkf-physics generate_sdf_contacts,
 * sdf_collision.rs:188-225) for many rigid bodies against the scene's field:
 * each sample point of the body's shape (CollisionShape::sample_points,
 * sdf_collision.rs:79-148) is placed at position + rotation * point; a point
 * whose distance is below the body's threshold yields a contact with the
 * central-difference normal (gradient_central_diff, eps 0.01) and
 * penetration threshold - distance. A body's contacts are sorted deepest
 * first (ties keep sample order). */
#define RKFM_BODY_SPHERE 0u  /* dims[0] = radius                             */
#define RKFM_BODY_BOX 1u     /* dims = half extents                           */
#define RKFM_BODY_CAPSULE 2u /* dims[0] = half height, dims[1] = radius       */
#define RKFM_MAX_CONTACTS 16u

typedef struct rkfm_body {
    float position[3];
    uint32_t shape;          /* RKFM_BODY_*                                   */
    float rotation[4];       /* unit quaternion x, y, z, w                    */
    float dims[3];
    float threshold;         /* contact_threshold                             */
} rkfm_body;

typedef struct rkfm_contact {
    float position[3];
    float normal[3];
    float penetration;
    uint32_t sample;         /* index of the shape's sample point             */
} rkfm_contact;

/* bodies, contacts (RKFM_MAX_CONTACTS per body) and counts (one per body)
 * are device pointers. */
int rkfm_sparse_contacts(rkfm_ctx *ctx, const rkfm_sparse_scene *scene,
                         const rkfm_body *bodies, uint32_t n,
                         rkfm_contact *contacts, uint32_t *counts,
                         void *hip_stream);

/* Free-space structures used by the DDA: free_idx has one u32 per L2 cell
 * (0xFFFFFFFF = every brick of the cell is free, else a row of free_bits);
 * free_bits holds 8 u64 per row, bit (x + 8y + 64z) set = that brick is not
 * free; oct has one u32 per L2 cell holding 8 octant extents of 4 bits
 * (octant o: bit a of o set = axis a negative), each the side in cells
 * (capped at 15) of the largest cube of entirely free cells with this cell at
 * its corner, extending in the octant's directions (0 = cell not free);
 * brick_ext holds per row 8 octant planes of 512 brick extents, 4 bits each
 * (byte (row * 8 + o) * 256 + lb / 2, low nibble for even lb): the side in
 * bricks (capped at 15) of a cube of free bricks with that brick at its
 * corner (0 = brick not free). *n_rows receives the row count (free_bits
 * needs 8 * n_rows words, brick_ext 2048 * n_rows bytes). */
int rkfm_sparse_download_free(rkfm_ctx *ctx, const rkfm_sparse_scene *scene,
                              uint32_t *free_idx, uint64_t *free_bits,
                              uint32_t *oct, uint8_t *brick_ext,
                              uint64_t *n_rows);

/* CSG / sculpt edits of a resident sparse scene: rkf-edit's EditOp
 * (crates/rkf-edit/src/edit_op.rs:15-28, enums types.rs:19-124) applied as
 * apply_edit_cpu (cpu_apply.rs:28-188) to the allocated bricks that
 * find_affected_bricks (edit_op.rs:97-140) selects, the scene's global brick
 * grid standing in for one object's brick map. */
#define RKFM_EDIT_CSG_UNION 0u
#define RKFM_EDIT_CSG_SUBTRACT 1u
#define RKFM_EDIT_CSG_INTERSECT 2u
#define RKFM_EDIT_SMOOTH_UNION 3u
#define RKFM_EDIT_SMOOTH_SUBTRACT 4u
#define RKFM_EDIT_SMOOTH 5u
#define RKFM_EDIT_FLATTEN 6u
#define RKFM_EDIT_PAINT 7u
#define RKFM_EDIT_COLOR_PAINT 8u
#define RKFM_SHAPE_SPHERE 0u
#define RKFM_SHAPE_BOX 1u
#define RKFM_SHAPE_CAPSULE 2u
#define RKFM_SHAPE_CYLINDER 3u
#define RKFM_SHAPE_TORUS 4u
#define RKFM_SHAPE_PLANE 5u
#define RKFM_FALLOFF_LINEAR 0u
#define RKFM_FALLOFF_SMOOTH 1u
#define RKFM_FALLOFF_SHARP 2u

typedef struct rkfm_edit {
    float position[3];      /* world (scene) space                          */
    uint32_t edit_type;     /* RKFM_EDIT_*                                  */
    float rotation[4];      /* unit quaternion x, y, z, w                   */
    float dimensions[3];    /* shape parameters (types.rs:71-90)            */
    uint32_t shape_type;    /* RKFM_SHAPE_*                                 */
    float strength;
    float blend_k;
    uint32_t falloff;       /* RKFM_FALLOFF_*                               */
    uint32_t material_id;   /* stored as its low 6 bits (voxel.rs:35)       */
} rkfm_edit;

/* Applies the edits in order (each voxel sees the previous edits' result)
 * in one launch. When n_modified is not NULL the call waits for the device
 * and reports the slots of bricks some edit changed by more than 0.1 vs in
 * some voxel (or painted), sorted: *n_modified = their count, the first
 * max_modified are copied to modified (may be NULL). Edits neither allocate
 * nor free bricks. */
int rkfm_sparse_edit(rkfm_ctx *ctx, rkfm_sparse_scene *scene,
                     const rkfm_edit *ops, uint32_t n_ops, uint32_t *modified,
                     uint64_t max_modified, uint64_t *n_modified,
                     void *hip_stream);

/* .rkf v2 object assets (crates/rkf-core/src/asset_file.rs): read on the
 * host (header, LOD table, LZ4 data blocks), written (save_object,
 * :380-460, LODs stored coarsest first), and streamed into resident sparse
 * scenes. Host-only functions: no device needed except for from_rkf. */
typedef struct rkfm_rkf_header {  /* RkfFileHeader, asset_file.rs:84-108 */
    char magic[4];                 /* "RKF2"                                 */
    uint32_t version;              /* 2                                      */
    uint32_t lod_count;
    uint32_t material_count;
    float aabb_min[3];
    float aabb_max[3];
    uint32_t analytical_type;      /* 0 none, 1 sphere .. 6 plane            */
    float analytical_params[4];
    uint16_t material_ids[16];
    uint8_t _reserved[36];
} rkfm_rkf_header;

typedef struct rkfm_rkf_lod {     /* LodFileEntry, asset_file.rs:118-136     */
    float voxel_size;
    uint32_t brick_count;
    uint32_t brick_dims[3];
    uint32_t _pad;
    uint64_t data_offset;
    uint32_t compressed_size;
    uint32_t uncompressed_size;
} rkfm_rkf_lod;

typedef struct rkfm_rkf_lod_src { /* SaveLodLevel, asset_file.rs:152-170    */
    float voxel_size;
    uint32_t dims[3];
    const uint32_t *map;           /* slots or RKF_EMPTY_SLOT; renumbered in   */
                                   /* order of first appearance               */
    const rkf_voxel_sample *bricks;/* brick_count x 512, in that order         */
    uint32_t brick_count;
    uint32_t _pad;
} rkfm_rkf_lod_src;

/* RKFM_ERR_FORMAT: bad magic / version / sizes; copies up to max_lods
 * entries of the LOD table. */
int rkfm_rkf_read_header(const void *file, uint64_t len, rkfm_rkf_header *hdr,
                         rkfm_rkf_lod *lods, uint32_t max_lods);
/* Decodes LOD `lod` (index into the stored, coarsest-first table): map gets
 * dims product u32 (bricks numbered 0..n-1), bricks n x 512 samples. Either
 * may be NULL. RKFM_ERR_OUT_OF_RANGE for a missing LOD, RKFM_ERR_FORMAT for
 * corrupt data. */
int rkfm_rkf_decode_lod(const void *file, uint64_t len, uint32_t lod,
                        uint32_t *map, rkf_voxel_sample *bricks);
/* Writes an asset (magic, version and LOD count are filled in; aabb,
 * analytical bound and materials are taken from hdr). out == NULL: *len
 * receives the size. */
int rkfm_rkf_encode(const rkfm_rkf_header *hdr, const rkfm_rkf_lod_src *lods,
                    uint32_t n_lods, void *out, uint64_t cap, uint64_t *len);
/* Decodes LOD `lod` and makes it a resident sparse scene whose grid corner
 * sits at `origin` (NULL: -dims * brick / 2, the centred grid voxelize_sdf
 * builds). */
int rkfm_sparse_from_rkf(rkfm_ctx *ctx, const void *file, uint64_t len,
                         uint32_t lod, const float origin[3],
                         rkfm_sparse_scene **out);

/* ------------------------------------------------------------------------ */
/* Batched launch: many uneven ray/query batches, over any resident sparse   */
/* scenes, in one kernel launch (capturable in a HIP graph). Each batch is   */
/* computed exactly as its single-batch entry point computes it.            */
/* ------------------------------------------------------------------------ */
#define RKFM_BATCH_VIEW 0u   /* camera view: G-buffer (+ shadow plane)        */
#define RKFM_BATCH_SHADOW 1u /* soft-shadow rays from given origins          */
#define RKFM_BATCH_POINTS 2u /* point queries                                */
#define RKFM_BATCH_SPHERES 3u /* sphere casts (rkfm_sparse_sphere_cast)      */

typedef struct rkfm_batch_item {
    uint32_t kind;                    /* RKFM_BATCH_*                         */
    uint32_t n;                       /* shadow rays / points                 */
    const rkfm_sparse_scene *scene;
    /* view */
    rkf_camera_uniforms camera;
    rkf_scene_uniforms su;
    rkfm_scene_ext ext;               /* ext.light_dir etc. also used by shadow batches;
                                         march_flags: views may set RKFM_MARCH_TILE_CULL
                                         (the tile intervals are computed in the same
                                         launch sequence), other kinds and the request
                                         server take 0 */
    rkfm_gbuffer out;
    /* shadow: origins (3 floats each, device) -> visibility (device) */
    const float *origins;
    float *visibility;
    /* points: xyz (3 floats each, device) -> results (device) */
    const float *xyz;
    rkfm_point_result *results;
    /* sphere casts: casts (device) -> cast results (device) */
    const rkfm_sphere_cast *casts;
    rkfm_cast_result *cast_results;
} rkfm_batch_item;

#define RKFM_SCHEDULE_GRID 0u       /* one wavefront per work unit           */
#define RKFM_SCHEDULE_PERSISTENT 1u /* persistent wavefronts pull work units */
/* One wavefront per work unit with the per-ray code of the single-view
 * kernels (each ray keeps the slots of its current brick neighbourhood:
 * fewer dependent loads per step, more memory instructions): for batch sets
 * whose time is set by their slowest rays rather than by their total work. */
#define RKFM_SCHEDULE_GRID_LOWLAT 2u

/* Launch shape: views are cut into tile_w x tile_h pixel packets
 * (tile_w * tile_h == 64, one packet per wavefront); shadow and point batches
 * into packets of 64. waves_per_cu sizes the persistent grid. */
typedef struct rkfm_batch_tuning {
    uint32_t schedule;
    uint32_t tile_w, tile_h;
    uint32_t waves_per_cu;
} rkfm_batch_tuning;

/* Runs the batches (NULL tuning = the autotuned or default shape). */
int rkfm_batch_run(rkfm_ctx *ctx, const rkfm_batch_item *items, uint32_t n,
                   const rkfm_batch_tuning *tuning, void *hip_stream);
/* Times every candidate launch shape on these batches (reps launches each)
 * and writes the fastest to *best; *best_ms receives its time per launch. */
int rkfm_batch_autotune(rkfm_ctx *ctx, const rkfm_batch_item *items,
                        uint32_t n, uint32_t reps, rkfm_batch_tuning *best,
                        float *best_ms);

/* A batch set captured once into a HIP graph (tile intervals of culled views,
 * work-counter reset, launch); replays re-run the same batches (their input
 * buffers may change between replays; outputs are rewritten). A graph uses
 * its context's scenes: destroy it before the context. Replays may run on any
 * stream; destroy waits for the last replay (an event recorded by every
 * launch) before it frees the graph's buffers. */
typedef struct rkfm_batch_graph rkfm_batch_graph;
int rkfm_batch_graph_create(rkfm_ctx *ctx, const rkfm_batch_item *items,
                            uint32_t n, const rkfm_batch_tuning *tuning,
                            rkfm_batch_graph **out);
int rkfm_batch_graph_launch(rkfm_batch_graph *g, void *hip_stream);
int rkfm_batch_graph_destroy(rkfm_batch_graph *g);
/* Rays + queries one run of the batches processes (views count primary
 * pixels; shadow rays of views are not known before the run). */
uint64_t rkfm_batch_work_items(const rkfm_batch_item *items, uint32_t n);

/* ------------------------------------------------------------------------ */
/* Request server: scenes stay resident, ray/query requests arrive from any  */
/* host thread, are admitted against a device-memory budget and run in      */
/* launch rounds that take every admitted request (continuous batching:     */
/* up to rounds_in_flight rounds overlap on the device, and requests that   */
/* arrive meanwhile join the next round as slots and memory free up). No    */
/* reference counterpart: the reference renders one view per frame on its   */
/* engine thread (engine_loop.rs:70-90) and answers spatial queries on the  */
/* CPU (rkf-editor/src/engine/query.rs:108-190).                           */
/* ------------------------------------------------------------------------ */
typedef struct rkfm_server rkfm_server;

#define RKFM_SERVE_BATCHED 0u     /* one batched launch per round (with the */
                                  /* default tuning, a round of one request */
                                  /* takes that request's plain launch)     */
#define RKFM_SERVE_PER_REQUEST 1u /* one plain launch per request, one at a */
                                  /* time (the per-request baseline)        */

typedef struct rkfm_server_config {
    uint64_t memory_budget;    /* device bytes for request inputs/outputs,   */
                               /* allocated once (0 = the context's budget,  */
                               /* else 1 GiB)                               */
    uint32_t max_requests;     /* admitted-request slots (0 = 1024)          */
    uint32_t rounds_in_flight; /* 1..4 (0 = 4); PER_REQUEST forces 1         */
    uint64_t max_round_work;   /* rays + queries per round (0 = 2^20; a      */
                               /* round always takes at least one request,  */
                               /* and at least the admitted queue's work /  */
                               /* rounds_in_flight; UINT64_MAX = no cap)    */
    uint32_t max_round_requests; /* requests per round (0 = 512)             */
    uint32_t policy;           /* RKFM_SERVE_*                               */
    rkfm_batch_tuning tuning;  /* batched launch shape (zero = 8x8 grid,     */
                               /* low-latency step code)                    */
    uint32_t max_pending;      /* submitted, not yet admitted (0 = 65536)    */
    uint32_t _reserved;
} rkfm_server_config;

/* Request flag: a view's G-buffer stays in the server's memory until
 * rkfm_server_release (read it with rkfm_server_view_planes); otherwise it is
 * copied to view_out's host planes. */
#define RKFM_REQ_KEEP_ON_DEVICE 1u

typedef struct rkfm_request {
    uint32_t kind;             /* RKFM_BATCH_VIEW / _SHADOW / _POINTS        */
    uint32_t scene;            /* index into the server's scene table        */
    uint32_t n;                /* shadow rays / points                       */
    uint32_t flags;            /* RKFM_REQ_*                                 */
    uint32_t width, height;    /* views                                      */
    rkf_camera_uniforms camera;
    rkf_scene_uniforms su;
    rkfm_scene_ext ext;        /* views: shadow plane written iff            */
                               /* ext.shadow_enable; shadow batches: light   */
    const float *in;           /* host: shadow origins / points (3 floats) / */
                               /* rkfm_sphere_cast records (8 floats)        */
    void *out;                 /* host: float visibility[n] /                */
                               /* rkfm_point_result[n] / rkfm_cast_result[n] */
    rkfm_gbuffer view_out;     /* views: host planes (NULL plane = skipped;  */
                               /* width/height ignored)                     */
    uint64_t tag;              /* caller's value, returned in the status     */
} rkfm_request;
/* Host buffers must stay valid until the request is done; memory from
 * rkfm_host_alloc (pinned) keeps the copies asynchronous. */

#define RKFM_REQ_QUEUED 0u   /* waiting for memory / a slot                  */
#define RKFM_REQ_ADMITTED 1u /* holds its memory, waiting for a round        */
#define RKFM_REQ_RUNNING 2u  /* in a launched round                          */
#define RKFM_REQ_DONE 3u
#define RKFM_REQ_FAILED 4u

typedef struct rkfm_request_status {
    uint32_t state;            /* RKFM_REQ_*                                 */
    int32_t error;             /* status code if FAILED                      */
    uint64_t tag;
    uint64_t round;            /* launch round that ran it                   */
    uint32_t round_requests;   /* requests in that round                     */
    uint32_t _pad;
    double t_submit, t_admit, t_launch, t_done; /* seconds, rkfm_server_now  */
} rkfm_request_status;

typedef struct rkfm_server_stats {
    uint64_t submitted, completed, failed, rounds;
    uint64_t work_items;       /* rays + queries of completed requests       */
    uint64_t arena_bytes, arena_in_use, arena_peak;
    uint32_t max_round_requests_seen, queued, admitted, running;
    double gpu_seconds;        /* sum of round durations (device events)     */
} rkfm_server_stats;

/* The scenes must stay resident (not destroyed) while the server runs. The
 * server owns a worker thread and its own streams; the context must outlive
 * it. */
int rkfm_server_create(rkfm_ctx *ctx, const rkfm_sparse_scene *const *scenes,
                       uint32_t n_scenes, const rkfm_server_config *cfg,
                       rkfm_server **out);
/* Thread-safe. RKFM_ERR_OOM: the request can never fit the budget;
 * RKFM_ERR_BUSY: max_pending requests are already queued. */
int rkfm_server_submit(rkfm_server *s, const rkfm_request *req, uint64_t *id);
int rkfm_server_status(rkfm_server *s, uint64_t id, rkfm_request_status *st);
/* Waits until the request is DONE or FAILED; RKFM_ERR_NOT_READY after
 * timeout_s seconds (negative = no limit). */
int rkfm_server_wait(rkfm_server *s, uint64_t id, double timeout_s,
                     rkfm_request_status *st);
/* Device planes of a DONE view kept on the device. */
int rkfm_server_view_planes(rkfm_server *s, uint64_t id, rkfm_gbuffer *planes);
/* Forgets a DONE or FAILED request and frees what it still holds. Every
 * request is released once its results are consumed. */
int rkfm_server_release(rkfm_server *s, uint64_t id);
int rkfm_server_stats_get(rkfm_server *s, rkfm_server_stats *st);
double rkfm_server_now(rkfm_server *s);
/* Finishes the launched rounds, fails what is still queued, joins. */
int rkfm_server_destroy(rkfm_server *s);

/* Replays an arrival trace from a native client thread: request i is
 * submitted at arrival_s[i] seconds after the call (arrival_s ascending),
 * waited for and released; latency_s[i] = completion - arrival (NaN if it
 * failed). *wall_s receives the time from the call to the last completion. */
int rkfm_server_replay(rkfm_server *s, const rkfm_request *reqs,
                       const double *arrival_s, uint32_t n, double *latency_s,
                       double *wall_s);

/* ------------------------------------------------------------------------ */
/* Device memory and timing helpers (the C ABI takes device pointers)        */
/* ------------------------------------------------------------------------ */
/* Pinned host memory (asynchronous copies to and from the device). */
int rkfm_host_alloc(uint64_t bytes, void **host_ptr);
int rkfm_host_free(void *host_ptr);
int rkfm_alloc(rkfm_ctx *ctx, uint64_t bytes, void **dev_ptr);
int rkfm_free(rkfm_ctx *ctx, void *dev_ptr);
int rkfm_memcpy_h2d(rkfm_ctx *ctx, void *dst_dev, const void *src_host,
                    uint64_t bytes);
int rkfm_memcpy_d2h(rkfm_ctx *ctx, void *dst_host, const void *src_dev,
                    uint64_t bytes);
int rkfm_memset(rkfm_ctx *ctx, void *dst_dev, int value, uint64_t bytes);
/* Plain copy kernel (bytes a multiple of 16) on the context stream; its
 * measured HBM traffic calibrates profiler byte counters. */
int rkfm_device_copy(rkfm_ctx *ctx, void *dst_dev, const void *src_dev,
                     uint64_t bytes);
/* Scattered-read kernel (n <= src_bytes / 128 lanes, src_bytes / 128 a power
 * of two): lane i reads the first 2 bytes of 128-byte line (i * 0x9E3779B1)
 * mod (src_bytes / 128) of src (every line at most once) into dst[i] (2 B per
 * lane). Its measured traffic calibrates profiler byte counters for
 * scattered 2-byte loads like the march's distance gathers. */
int rkfm_device_gather16(rkfm_ctx *ctx, void *dst_dev, const void *src_dev,
                         uint64_t src_bytes, uint64_t n);
/* HIP events on the context stream around the work enqueued in between. */
int rkfm_timer_begin(rkfm_ctx *ctx);
int rkfm_timer_end(rkfm_ctx *ctx, float *ms);

#ifdef __cplusplus
} /* extern "C" */
#endif

/* Layout checks (the reference asserts the same sizes in its unit tests:
 * gpu_object.rs:272, camera.rs:220, gpu_scene.rs:380). */
#ifdef __cplusplus
static_assert(sizeof(rkf_voxel_sample) == 8, "VoxelSample is 8 bytes");
static_assert(sizeof(rkf_gpu_object) == 256, "GpuObject is 256 bytes");
static_assert(sizeof(rkf_camera_uniforms) == 144, "CameraUniforms is 144 bytes");
static_assert(sizeof(rkf_scene_uniforms) == 16, "SceneUniforms is 16 bytes");
static_assert(sizeof(rkf_coarse_field_uniforms) == 48, "CoarseFieldUniforms is 48 bytes");
static_assert(sizeof(rkf_bvh_node) == 32, "BvhNode is 32 bytes");
static_assert(sizeof(rkfm_batch_item) == 328, "batch item record");
static_assert(sizeof(rkfm_request) == 312, "request record");
static_assert(sizeof(rkfm_server_config) == 56, "server config record");
static_assert(sizeof(rkfm_request_status) == 64, "request status record");
static_assert(sizeof(rkfm_server_stats) == 88, "server stats record");
static_assert(sizeof(rkfm_edit) == 64, "edit record");
static_assert(sizeof(rkfm_body) == 48, "body record");
static_assert(sizeof(rkfm_rkf_header) == 128, "RkfFileHeader is 128 bytes");
static_assert(sizeof(rkfm_rkf_lod) == 40, "LodFileEntry is 40 bytes");
static_assert(sizeof(rkfm_contact) == 32, "contact record");
static_assert(sizeof(rkfm_sphere_cast) == 32, "sphere cast record");
static_assert(sizeof(rkfm_cast_result) == 32, "cast result record");
#else
_Static_assert(sizeof(rkf_voxel_sample) == 8, "VoxelSample is 8 bytes");
_Static_assert(sizeof(rkf_gpu_object) == 256, "GpuObject is 256 bytes");
_Static_assert(sizeof(rkf_camera_uniforms) == 144, "CameraUniforms is 144 bytes");
_Static_assert(sizeof(rkf_scene_uniforms) == 16, "SceneUniforms is 16 bytes");
_Static_assert(sizeof(rkf_coarse_field_uniforms) == 48, "CoarseFieldUniforms is 48 bytes");
_Static_assert(sizeof(rkf_bvh_node) == 32, "BvhNode is 32 bytes");
_Static_assert(sizeof(rkfm_batch_item) == 328, "batch item record");
_Static_assert(sizeof(rkfm_request) == 312, "request record");
_Static_assert(sizeof(rkfm_server_config) == 56, "server config record");
_Static_assert(sizeof(rkfm_request_status) == 64, "request status record");
_Static_assert(sizeof(rkfm_server_stats) == 88, "server stats record");
_Static_assert(sizeof(rkfm_edit) == 64, "edit record");
_Static_assert(sizeof(rkfm_body) == 48, "body record");
_Static_assert(sizeof(rkfm_rkf_header) == 128, "RkfFileHeader is 128 bytes");
_Static_assert(sizeof(rkfm_rkf_lod) == 40, "LodFileEntry is 40 bytes");
_Static_assert(sizeof(rkfm_contact) == 32, "contact record");
_Static_assert(sizeof(rkfm_sphere_cast) == 32, "sphere cast record");
_Static_assert(sizeof(rkfm_cast_result) == 32, "cast result record");
#endif

#endif /* RKF_MARCH_H */

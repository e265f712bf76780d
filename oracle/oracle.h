/*
 * oracle.h — CPU oracle for the rkifield ray-march hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference's
 * algorithm (Rust crates/rkf-core + WGSL crates/rkf-render/shaders/ray_march.wgsl),
 * used by tests/, the smoke check and the CPU-baseline leg of bench.py as the
 * checker. The product library (rkifield_amd/) never links, loads or calls it.
 *
 * Every function cites the reference file:line it follows. The march is built
 * twice from trace_impl.h: an FP64 oracle (depth / normal tolerances) and an
 * FP32 twin with the kernels' exact operation order (bit-exact integer fields).
 *
 * Pins (what fixes each function; tests/test_oracle.py, tests/test_oracle_pins.py,
 * tests/test_rkf_asset.py):
 *   f16 encode/decode          voxel.rs tests (1.5, -0.5, f16::MAX, +inf) and an
 *                              exhaustive comparison with IEEE binary16 (numpy)
 *   brick_index / flatten      brick.rs / brick_map.rs test vectors (0,7,56,448,511; 59)
 *   pool slot order            brick_pool.rs tests (0,1,2,3)
 *   sample_brick_trilinear     sampling.rs:88-292 test vectors
 *   voxelize_sdf, .rkf reader  golden fixture scenes/test_cross.rkf (224 bricks), bit-exact
 *                              through an LZ4 decoder independent of the product's;
 *                              asset_file_tests.rs round trips
 *   sample_voxelized (global)  brute force: equals sample_brick_trilinear at points
 *                              more than one voxel from brick faces; constant field
 *   analytic primitives        each of the six read at points offset along its normal
 *   transforms                 rotated, non-uniformly scaled sphere: hits on the
 *                              ray-ellipsoid intersection, closed-form ellipsoid normal
 *   march                      voxelised sphere depth -> ray-sphere intersection
 *                              within O(vs); all-EMPTY map misses; FP64 vs FP32
 *                              agreement; a planted mistake is caught
 *   bisection                  planar crossing, t known to step * 2^-8
 *   normals                    closed form: sphere normal (p-c)/r within O(vs)
 *   motion                     camera translated along its right axis: pixel shift
 *   smin                       smin(a, a, k) = a - k/4; the |a - b| >= k branch
 *   soft shadow                unoccluded ray 1 exactly; through a voxelised sphere 0;
 *                              past a plane at known height: the min(k d / t) form
 *   point query, sphere cast   ground plane: distance y, inside y < 0, gradient
 *                              (0,1,0); exact contact distance and normal
 *   DDA extents, tile intervals  brute force and soundness by dense sampling
 *   edits, contacts            cpu_apply.rs / sdf_collision.rs tests, closed forms
 *   coarse field / tile cull   closed form AABB distances; projection of known boxes
 *
 * Parity unpinned (nothing but the restatement fixes them): the hardware
 * trilinear coarse-field sampler (restated as a software sampler); glam's
 * Mat4::inverse (supplied by the host) and Quat * Vec3 (scalar formula
 * restated); naga's lowering of normalize, mix and unpack2x16float; the BVH
 * topology of Bvh::build's SAH; lz4_flex's exact compressed bytes (any valid
 * block decodes).
 */
#ifndef RKF_ORACLE_H
#define RKF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/rkf_march.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- binary16 (the `half` crate's f16::from_f32 / to_f32: IEEE RNE) ---- */
uint16_t orc_f32_to_f16(float f);
float orc_f16_to_f32(uint16_t h);
void orc_f32_to_f16_array(const float *in, uint16_t *out, uint64_t n);
void orc_f16_to_f32_array(const uint16_t *in, float *out, uint64_t n);

/* ---- VoxelSample (voxel.rs:32-90) ---- */
rkf_voxel_sample orc_voxel_new(float distance, uint16_t material, uint8_t blend);
rkf_voxel_sample orc_voxel_new_blended(float distance, uint16_t material,
                                       uint8_t secondary, uint8_t blend);
float orc_voxel_distance(rkf_voxel_sample v);
uint32_t orc_voxel_material(rkf_voxel_sample v);
uint32_t orc_voxel_secondary(rkf_voxel_sample v);
uint32_t orc_voxel_blend(rkf_voxel_sample v);

/* ---- index math (brick.rs:28-42, brick_map.rs:80-85) ---- */
int64_t orc_brick_index(uint32_t x, uint32_t y, uint32_t z); /* -1 if out of range */
int64_t orc_map_flatten(const uint32_t dims[3], uint32_t bx, uint32_t by,
                        uint32_t bz); /* -1 = None */

/* ---- Pool<T> slot order (brick_pool.rs:37-47,119-129) ---- */
typedef struct orc_pool {
    uint32_t capacity;
    uint32_t *free_list; /* stack, top = last */
    uint32_t n_free;
} orc_pool;
int orc_pool_init(orc_pool *p, uint32_t capacity);
void orc_pool_free(orc_pool *p);
int64_t orc_pool_allocate(orc_pool *p);
int orc_pool_allocate_range(orc_pool *p, uint32_t count, uint32_t *slots);

/* ---- CPU sampling (sampling.rs:23-81) on one brick of 512 voxels ---- */
float orc_sample_brick_trilinear(const rkf_voxel_sample *brick, const float local_pos[3]);
uint32_t orc_sample_brick_nearest_material(const rkf_voxel_sample *brick,
                                           const float local_pos[3]);

/* ---- analytic SDFs (sdf.rs:12-51, voxelize_object.rs:439-474) ---- */
float orc_sphere_sdf(const float c[3], float r, const float p[3]);
float orc_box_sdf(const float half_ext[3], const float p[3]);
float orc_capsule_sdf(const float a[3], const float b[3], float r, const float p[3]);
float orc_smin(float a, float b, float k);
/* gen_test_rkf.rs:22-25 variant with k.max(1e-6) */
float orc_smooth_min_cross(float a, float b, float k);
/* fBm terrain height of a type-3 primitive (without the distance scale). */
float orc_terrain_height(const rkfm_prim *pr, float x, float z);
float orc_prim_eval(const rkfm_prim *pr, const float p[3]);
float orc_scene_field(const rkfm_prim *prims, uint32_t n, float k, const float p[3]);

/* ---- voxelize_sdf (voxelize_object.rs:53-182) ----
 * sdf(p, user) -> distance, writes material to *mat. The grid is centred at
 * the local origin. dims: out. map: dims product entries (EMPTY or slot).
 * bricks: 512 voxels per allocated brick, in slot order starting at
 * first_slot (slots are consecutive because a fresh Pool pops 0,1,2...).
 * Call with map == NULL to get dims and brick count only. */
typedef float (*orc_sdf_fn)(const float p[3], void *user, uint16_t *mat);
int orc_voxelize_sdf(orc_sdf_fn sdf, void *user, const float aabb_min[3],
                     const float aabb_max[3], float voxel_size,
                     uint32_t first_slot, uint32_t dims[3], uint32_t *map,
                     rkf_voxel_sample *bricks, uint64_t *n_bricks);
/* voxelize_to_geometry's classification (INTERIOR / EMPTY / stored bricks,
 * voxelize_object.rs:217-385) over voxelize_sdf's grid and voxel words. */
int orc_voxelize_geometry(orc_sdf_fn sdf, void *user, const float aabb_min[3], const float aabb_max[3],
                          float voxel_size, uint32_t first_slot, uint32_t dims[3], uint32_t *map,
                          rkf_voxel_sample *bricks, uint64_t *n_bricks);
float orc_prim_sdf(const float p[3], void *user, uint16_t *mat);
float orc_test_cross_sdf(const float p[3], void *user, uint16_t *mat);
void orc_test_cross_aabb(float mn[3], float mx[3]);

/* ---- .rkf v2 loader (asset_file.rs:74-150,386-590) + LZ4 block format ---- */
int orc_lz4_decompress_block(const uint8_t *src, uint64_t src_len, uint8_t *dst,
                             uint64_t dst_cap, uint64_t *out_len);
typedef struct orc_rkf_header {
    uint32_t version, lod_count, material_count;
    float aabb_min[3], aabb_max[3];
    uint16_t material_ids[16];
    float voxel_size;
    uint32_t brick_count;
    uint32_t dims[3];
    uint64_t data_offset;
    uint32_t compressed_size, uncompressed_size;
} orc_rkf_header;
/* Parses the header + LOD 0 entry; then decodes LOD 0 into map (dims product
 * u32) and bricks (brick_count*512 voxels). map/bricks may be NULL for a
 * header-only parse. Returns 0 on success, negative on format error. */
int orc_rkf_load(const uint8_t *file, uint64_t len, orc_rkf_header *hdr,
                 uint32_t *map, rkf_voxel_sample *bricks);

/* ---- producers: coarse field and tile cull ---- */
/* CoarseField::populate (coarse_field.rs:240-266), f16 cells x fastest. */
void orc_coarse_populate(const float origin[3], const uint32_t dims[3],
                         float voxel_size, const rkf_gpu_object *objects,
                         uint32_t n_objects, uint16_t *cells);
/* tile_object_cull.wgsl:117-204, ordered by object index, first 32 kept. */
void orc_tile_cull(const rkf_camera_uniforms *cam,
                   const rkf_scene_uniforms *su,
                   const rkf_gpu_object *objects, uint32_t width,
                   uint32_t height, uint32_t *idx, uint32_t *counts);

/* ---- march ---- */
typedef struct orc_ref_scene {
    const rkf_voxel_sample *pool;
    uint64_t n_pool_voxels;
    const uint32_t *maps;
    uint64_t n_maps;
    const rkf_gpu_object *objects;
    const rkf_bvh_node *bvh;
    uint32_t n_bvh;
    const uint32_t *tile_idx;
    const uint32_t *tile_counts;
    uint32_t tiles_x, tiles_y;
    const uint16_t *coarse;                    /* NULL: field absent */
    rkf_coarse_field_uniforms coarse_u;
    rkf_camera_uniforms cam;
    rkf_scene_uniforms su;
    rkfm_scene_ext ext;
    uint32_t width, height;
} orc_ref_scene;

typedef struct orc_sparse_scene {
    float origin[3];
    float voxel_size;
    uint32_t cells[3];
    const uint32_t *l2;
    const uint32_t *l1;
    const uint16_t *dist;
    const uint8_t *mat;
    const uint8_t *free_l2;    /* per L2 cell: 1 = every brick of it is free */
    const uint32_t *oct;       /* per L2 cell: 8 octant extents (4 bits each) */
    const uint32_t *bext;      /* per L2 cell x 512 bricks: 8 brick octant
                                  extents (4 bits each); used in cells that
                                  are not entirely free */
    const uint16_t *aux;       /* per slot x 512: secondary | blend << 8, or
                                  NULL (both 0) */
} orc_sparse_scene;

/* Unpacked per-ray result. For the FP32 twin all values are exactly the
 * float results widened to double. */
/* Branches a ray's evaluations took (orc_ray_result.flags): a stepping
 * sample read an INTERIOR corner (-2 vs), a gradient tap read one (-8 vs),
 * a geometry-AABB early-out returned, the overstep bisection ran, the hit
 * object is analytic. */
#define ORC_FLAG_INTERIOR_STEP 1u
#define ORC_FLAG_INTERIOR_GRAD 2u
#define ORC_FLAG_GEOMETRY_EARLY_OUT 4u
#define ORC_FLAG_BISECT 8u
#define ORC_FLAG_ANALYTIC_HIT 16u

/* Hash of a march's discrete decisions, one code per loop iteration (free
 * skip, sample, near sample + candidate object) and per bisection halving:
 * FP32 and FP64 rays with equal hashes took the same path, so their depth
 * and normal differ by rounding alone. */
#define ORC_PATH(h, code) ((h) = (h) * 31u + (uint32_t)(code))

/* u32(f32(x)) with WGSL's saturating conversion. */
static inline uint32_t orc_u32_via_f32(uint32_t x) {
    float f = (float)x;
    return f >= 4294967295.0f ? 0xFFFFFFFFu : (uint32_t)f;
}

typedef struct orc_ray_result {
    int32_t hit;
    uint32_t flags;            /* ORC_FLAG_* */
    double t;
    double pos[3];
    double normal[3];
    double motion[2];
    double sdf_at_hit;
    uint32_t packed_material;
    uint32_t material_id;
    uint32_t obj_idx;
    uint32_t object_id;
    uint32_t brick_slot;
    uint32_t voxel_index;
    uint32_t cell_index;
    uint32_t steps;
    uint32_t evals;
    uint32_t shadow_steps;
    uint32_t shadow_evals;
    uint32_t path;             /* ORC_PATH hash of the primary march */
    double shadow;
} orc_ray_result;

/* Trace pixels [p0, p1) of the image (row-major). prec: 32 = FP32 twin,
 * 64 = FP64 oracle. */
void orc_ref_trace(const orc_ref_scene *s, int prec, uint64_t p0, uint64_t p1,
                   orc_ray_result *out);
/* Sparse scene: free_l2 is derived by orc_sparse_free_l2 from l2/l1 (a
 * brick b is free when every brick of {b..b+1}^3, clamped to the grid, is
 * EMPTY; that test is evaluated on the fly). */
uint64_t orc_sparse_n_bricks_global(const orc_sparse_scene *s);
void orc_sparse_free_l2(const orc_sparse_scene *s, uint8_t *free_l2);
/* Octant extents from free_l2: for octant o (bit a of o set = axis a runs
 * negative) E_o(C) = 0 if C is not free, else the side (capped at 15) of the
 * largest cube of free cells with C at its corner extending in o's
 * directions, out-of-grid cells counting as free. Computed by one sweep per
 * octant of E = min(15, 1 + min over the 7 neighbours further along o). */
void orc_sparse_oct(const orc_sparse_scene *s, const uint8_t *free_l2, uint32_t *oct);
/* Brick octant extents for every brick of every cell that is not entirely
 * free (out[cell * 512 + lb], 0 elsewhere): E_o(b) = 0 if brick b is not
 * free, else min(15, 1 + min over the 7 neighbours further along o), where a
 * neighbour outside the brick grid counts 15 and a neighbour n in an entirely
 * free cell C counts min(15, distance from n to the far side of the free
 * cube of cells C + [0, E_o(C)) on each axis). One sweep per octant. */
void orc_sparse_bext(const orc_sparse_scene *s, const uint8_t *free_l2, const uint32_t *oct, uint32_t *out);
int orc_sparse_brick_free(const orc_sparse_scene *s, uint32_t bx, uint32_t by, uint32_t bz);
/* CSG / sculpt edits: for each op in order, find_affected_bricks
 * (edit_op.rs:97-140; the scene's brick grid with origin as the grid
 * corner) then apply_edit_cpu (cpu_apply.rs:28-188) brick by brick in z-y-x
 * order, voxel by voxel. dist / mat are the scene's planes, updated in place
 * (the scene's own pointers are only used for the index). modified[slot] is
 * set to 1 for every brick an op reports modified. */
void orc_sparse_edit(const orc_sparse_scene *s, uint16_t *dist, uint8_t *mat, const rkfm_edit *ops,
                     uint32_t n_ops, uint8_t *modified);
/* generate_sdf_contacts (sdf_collision.rs:188-225) per body over the FP32
 * sparse field: out has RKFM_MAX_CONTACTS records per body, deepest first
 * (stable), counts[i] = number of contacts of body i. */
void orc_sparse_contacts(const orc_sparse_scene *s, const rkfm_body *bodies, uint32_t n, rkfm_contact *out,
                         uint32_t *counts);
/* evaluate_shape / evaluate_falloff (cpu_apply.rs:231-281). */
float orc_edit_shape(uint32_t shape, const float dims[3], const float p[3]);
float orc_edit_falloff(uint32_t curve, float distance, float radius);
void orc_sparse_trace(const orc_sparse_scene *s, const rkf_camera_uniforms *cam,
                      const rkf_scene_uniforms *su, const rkfm_scene_ext *ext,
                      uint32_t width, uint32_t height, int prec, uint64_t p0,
                      uint64_t p1, orc_ray_result *out);
void orc_sparse_query(const orc_sparse_scene *s, const float *xyz, uint64_t n,
                      int prec, rkfm_point_result *out);
/* Tile intervals of RKFM_MARCH_TILE_CULL (same definition as
 * rkfm_sparse_tile_intervals): out gets (lo, hi) per 8x8 tile, row-major;
 * prec 32 values are the float results widened. When ext->march_flags has
 * RKFM_MARCH_TILE_CULL, orc_sparse_trace(_parallel) compute them at the
 * trace's precision and march every ray within its tile's interval. */
void orc_sparse_tiles(const orc_sparse_scene *s, const rkf_camera_uniforms *cam,
                      const rkf_scene_uniforms *su, uint32_t width, uint32_t height,
                      int prec, double *out);
/* Per 8x8 tile, the number of L2 cells (holding an L1 block) binned into it
 * by the FP32 tile binning: the key of the sparse kernel's tile launch order. */
void orc_sparse_tile_counts(const orc_sparse_scene *s, const rkf_camera_uniforms *cam,
                            const rkf_scene_uniforms *su, uint32_t width, uint32_t height, uint32_t *counts);
void orc_sparse_shadow(const orc_sparse_scene *s, const float *origins, uint64_t n, const float L[3],
                       float max_dist, float k, int prec, float *visibility);
/* Sphere casts as rkfm_sparse_sphere_cast (prec 32: FP32 twin, 64: FP64). */
void orc_sparse_cast(const orc_sparse_scene *s, const rkfm_sphere_cast *casts, uint64_t n, int prec,
                     rkfm_cast_result *out);
/* Sparse field value at a world point (stepping semantics). */
double orc_sparse_sample(const orc_sparse_scene *s, const float p[3], int prec);

/* evaluate_object (ray_march.wgsl:600-634) of object oi at world point p:
 * distance (FP32 twin widened, or FP64) and material. */
void orc_ref_eval_object(const orc_ref_scene *s, int prec, uint32_t oi, const double p[3], double *dist,
                         uint32_t *mat);

/* rkfm_sparse_from_objects' resampling of a reference scene (objects binned
 * by AABB + 3 bricks into L2 cells, field = min of evaluate_object, stored
 * bricks |field(centre)| < 1.8 bricks, slots in cell then z-y-x order), FP32.
 * Call with l2 == NULL to size (cells, n_blocks, n_bricks); then l2 holds
 * cells product u32, l1 n_blocks * 512, dist/aux n_bricks * 512 u16, mat
 * n_bricks * 512 u8 (aux may be NULL). */
int orc_sparse_from_objects(const orc_ref_scene *s, float vs, const float bmin[3], const float bmax[3],
                            uint32_t cells[3], uint32_t *n_blocks, uint64_t *n_bricks, uint32_t *l2, uint32_t *l1,
                            uint16_t *dist, uint8_t *mat, uint16_t *aux);

/* Pack unpacked results into the reference G-buffer formats. */
void orc_pack_gbuffer(const orc_ray_result *r, uint64_t n, float *position,
                      uint16_t *normal, uint32_t *material, float *motion);

/* CPU baseline: FP32 twin over all pixels with OpenMP dynamic scheduling in
 * chunks of 64 rays; returns the thread count used. */
int orc_ref_trace_parallel(const orc_ref_scene *s, orc_ray_result *out,
                           int threads);
int orc_sparse_trace_parallel(const orc_sparse_scene *s,
                              const rkf_camera_uniforms *cam,
                              const rkf_scene_uniforms *su,
                              const rkfm_scene_ext *ext, uint32_t width,
                              uint32_t height, uint64_t p0, uint64_t p1,
                              orc_ray_result *out, int threads);

#ifdef __cplusplus
}
#endif
#endif

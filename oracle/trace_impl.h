/*
 * trace_impl.h — the march, written once and compiled twice by trace.c:
 *   R = float  (SFX _f32): the FP32 twin. Same operation order as the HIP
 *                          kernels, no FMA contraction, IEEE div/sqrt.
 *   R = double (SFX _f64): the FP64 oracle (same algorithm, wider arithmetic).
 * Test infrastructure only (see oracle.h).
 *
 * Reference-layout march: crates/rkf-render/shaders/ray_march.wgsl
 *   sample_voxel_at            :270-288      sample_voxel_at_grad  :300-315
 *   sample_voxelized           :353-413      sample_voxelized_grad :463-498
 *   sample_voxelized_gradient  :504-513      material / blend      :540-596
 *   evaluate_object            :600-634      analytic primitives   :182-264
 *   ray_march_bvh / brute / tiled :641-985   normal :1003-1032  main :1036-1111
 *   sample_coarse_field        :228-235 (software trilinear, texel centres,
 *                                        clamp-to-edge, instead of the
 *                                        hardware sampler)
 * Sparse-scene march: the same field and step rules over the two-level index,
 * plus free-space DDA (defined in trace_sparse below).
 * Shadow rays: soft_shadow shade_common_shading.wgsl:9-33, origin
 * shade_pbr.wgsl:40, constants shade_common.wgsl:273-277.
 */
#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define F(name) CAT(name, SFX)

typedef struct { R x, y, z; } F(v3);
#define V3 F(v3)

static inline V3 F(mk)(R x, R y, R z) { V3 v = {x, y, z}; return v; }
static inline V3 F(add)(V3 a, V3 b) { return F(mk)(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 F(sub)(V3 a, V3 b) { return F(mk)(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 F(muls)(V3 a, R s) { return F(mk)(a.x * s, a.y * s, a.z * s); }
static inline R F(dot)(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline R F(length)(V3 a) { return SQRT(F(dot)(a, a)); }
static inline R F(minr)(R a, R b) { return FMIN(a, b); }
static inline R F(maxr)(R a, R b) { return FMAX(a, b); }
static inline R F(clampr)(R x, R lo, R hi) { return FMIN(FMAX(x, lo), hi); }
static inline R F(mix)(R a, R b, R t) { return a * ((R)1 - t) + b * t; }
static inline V3 F(normalize)(V3 a) { R l = F(length)(a); return F(mk)(a.x / l, a.y / l, a.z / l); }
static inline int32_t F(clampi)(int32_t x, int32_t lo, int32_t hi) { return x < lo ? lo : (x > hi ? hi : x); }

#define K_MAX_FLOAT ((R)3.402823e+38f)
#define K_MIN_STEP ((R)0.0005f)
#define K_NEAR ((R)0.02f)
#define K_COARSE_NEAR ((R)0.5f)
#define K_HARD_MAX 1024u
#define K_BISECT 8u

/* inverse_world * vec4(p, 1): column-major m[c*4+r], summed left to right. */
static inline V3 F(xform_point)(const float *m, V3 p) {
    return F(mk)((R)m[0] * p.x + (R)m[4] * p.y + (R)m[8] * p.z + (R)m[12],
                 (R)m[1] * p.x + (R)m[5] * p.y + (R)m[9] * p.z + (R)m[13],
                 (R)m[2] * p.x + (R)m[6] * p.y + (R)m[10] * p.z + (R)m[14]);
}
/* transpose(inverse_world) * vec4(g, 0) (ray_march.wgsl:1014). */
static inline V3 F(xform_grad)(const float *m, V3 g) {
    return F(mk)((R)m[0] * g.x + (R)m[1] * g.y + (R)m[2] * g.z + (R)m[3] * (R)0,
                 (R)m[4] * g.x + (R)m[5] * g.y + (R)m[6] * g.z + (R)m[7] * (R)0,
                 (R)m[8] * g.x + (R)m[9] * g.y + (R)m[10] * g.z + (R)m[11] * (R)0);
}

static inline R F(hit_eps)(const rkfm_scene_ext *ext, const rkf_scene_uniforms *su, R t) {
    return ext->hit_eps_mode == RKFM_EPS_RELATIVE ? (R)su->hit_threshold * t : (R)su->hit_threshold;
}

/* ================================================================== */
/* Reference layout                                                     */
/* ================================================================== */

typedef struct {
    const orc_ref_scene *s;
    uint32_t evals;
    uint32_t flags; /* ORC_FLAG_*: which branches this ray's evaluations took */
    uint32_t path;  /* ORC_PATH hash of the march's discrete decisions */
} F(rctx);

static inline uint32_t F(map_entry)(const orc_ref_scene *s, uint64_t i) {
    return i < s->n_maps ? s->maps[i] : RKF_EMPTY_SLOT;
}
static inline uint32_t F(pool_word0)(const orc_ref_scene *s, uint64_t i) {
    return i < s->n_pool_voxels ? s->pool[i].word0 : 0u;
}
static inline uint32_t F(pool_word1)(const orc_ref_scene *s, uint64_t i) {
    return i < s->n_pool_voxels ? s->pool[i].word1 : 0u;
}

/* :270-288 (grad = 0) and :300-315 (grad = 1) */
static R F(voxel_at)(const orc_ref_scene *s, uint32_t off, int32_t cx, int32_t cy, int32_t cz,
                     const uint32_t *dims, const int32_t *tot, R vs, int grad, uint32_t *flags) {
    cx = F(clampi)(cx, 0, tot[0] - 1);
    cy = F(clampi)(cy, 0, tot[1] - 1);
    cz = F(clampi)(cz, 0, tot[2] - 1);
    uint32_t bx = (uint32_t)(cx / 8), by = (uint32_t)(cy / 8), bz = (uint32_t)(cz / 8);
    uint32_t lx = (uint32_t)(cx % 8), ly = (uint32_t)(cy % 8), lz = (uint32_t)(cz % 8);
    uint32_t flat = bx + by * dims[0] + bz * dims[0] * dims[1];
    uint32_t slot = F(map_entry)(s, (uint64_t)off + flat);
    if (slot == RKF_EMPTY_SLOT) return vs * (R)8.0f;
    if (slot == RKF_INTERIOR_SLOT) {
        *flags |= grad ? ORC_FLAG_INTERIOR_GRAD : ORC_FLAG_INTERIOR_STEP;
        return grad ? -(vs * (R)8.0f) : -(vs * (R)2.0f);
    }
    uint64_t idx = (uint64_t)slot * 512u + lx + ly * 8u + lz * 64u;
    return (R)orc_f16_to_f32((uint16_t)(F(pool_word0)(s, idx) & 0xFFFFu));
}

/* :353-413 (grad = 0, with geometry-AABB early-out) and :463-498 (grad = 1). */
static R F(sample_voxelized)(F(rctx) *c, V3 lp, const rkf_gpu_object *o, int grad) {
    R vs = (R)o->voxel_size;
    R be = vs * (R)8.0f;
    const uint32_t *dims = o->brick_map_dims;
    V3 gs = F(mk)((R)dims[0] * be, (R)dims[1] * be, (R)dims[2] * be);
    V3 gp = F(add)(lp, F(muls)(gs, (R)0.5f));
    R lo = vs * (R)0.01f;
    V3 cl = F(mk)(F(clampr)(gp.x, lo, gs.x - lo), F(clampr)(gp.y, lo, gs.y - lo), F(clampr)(gp.z, lo, gs.z - lo));
    R outside = F(length)(F(sub)(gp, cl));
    if (outside > be * (R)2.0f) return outside;
    if (!grad) {
        V3 gmin = F(mk)((R)o->geometry_aabb_min[0], (R)o->geometry_aabb_min[1], (R)o->geometry_aabb_min[2]);
        V3 gmax = F(mk)((R)o->geometry_aabb_max[0], (R)o->geometry_aabb_max[1], (R)o->geometry_aabb_max[2]);
        if (gmax.x > gmin.x) {
            V3 gc = F(mk)(F(clampr)(lp.x, gmin.x, gmax.x), F(clampr)(lp.y, gmin.y, gmax.y),
                          F(clampr)(lp.z, gmin.z, gmax.z));
            R gd = F(length)(F(sub)(lp, gc));
            if (gd > be) {
                c->flags |= ORC_FLAG_GEOMETRY_EARLY_OUT;
                return gd;
            }
        }
    }
    V3 vc = F(mk)(cl.x / vs - (R)0.5f, cl.y / vs - (R)0.5f, cl.z / vs - (R)0.5f);
    R fx = FLOOR(vc.x), fy = FLOOR(vc.y), fz = FLOOR(vc.z);
    int32_t x0 = (int32_t)fx, y0 = (int32_t)fy, z0 = (int32_t)fz;
    R tx = vc.x - (R)x0, ty = vc.y - (R)y0, tz = vc.z - (R)z0;
    int32_t tot[3] = {(int32_t)dims[0] * 8, (int32_t)dims[1] * 8, (int32_t)dims[2] * 8};
    uint32_t off = o->brick_map_offset;
    const orc_ref_scene *s = c->s;
    if (!grad) c->evals++;
    R c000 = F(voxel_at)(s, off, x0, y0, z0, dims, tot, vs, grad, &c->flags);
    R c100 = F(voxel_at)(s, off, x0 + 1, y0, z0, dims, tot, vs, grad, &c->flags);
    R c010 = F(voxel_at)(s, off, x0, y0 + 1, z0, dims, tot, vs, grad, &c->flags);
    R c110 = F(voxel_at)(s, off, x0 + 1, y0 + 1, z0, dims, tot, vs, grad, &c->flags);
    R c001 = F(voxel_at)(s, off, x0, y0, z0 + 1, dims, tot, vs, grad, &c->flags);
    R c101 = F(voxel_at)(s, off, x0 + 1, y0, z0 + 1, dims, tot, vs, grad, &c->flags);
    R c011 = F(voxel_at)(s, off, x0, y0 + 1, z0 + 1, dims, tot, vs, grad, &c->flags);
    R c111 = F(voxel_at)(s, off, x0 + 1, y0 + 1, z0 + 1, dims, tot, vs, grad, &c->flags);
    R c00 = F(mix)(c000, c100, tx), c10 = F(mix)(c010, c110, tx);
    R c01 = F(mix)(c001, c101, tx), c11 = F(mix)(c011, c111, tx);
    R c0 = F(mix)(c00, c10, ty), c1 = F(mix)(c01, c11, ty);
    return F(mix)(c0, c1, tz) + outside;
}

/* Nearest-voxel lookup shared by :540-566 and :570-596. Returns the pool
 * index or UINT64_MAX; slot/voxel written for the debug plane. */
static uint64_t F(nearest_voxel)(const orc_ref_scene *s, V3 lp, const rkf_gpu_object *o,
                                 uint32_t *slot_out, uint32_t *vox_out) {
    R vs = (R)o->voxel_size;
    R be = vs * (R)8.0f;
    const uint32_t *dims = o->brick_map_dims;
    V3 gs = F(mk)((R)dims[0] * be, (R)dims[1] * be, (R)dims[2] * be);
    V3 gp = F(add)(lp, F(muls)(gs, (R)0.5f));
    *slot_out = RKF_EMPTY_SLOT;
    *vox_out = 0xFFFFFFFFu;
    if (gp.x < (R)0 || gp.y < (R)0 || gp.z < (R)0 || gp.x >= gs.x || gp.y >= gs.y || gp.z >= gs.z)
        return UINT64_MAX;
    int32_t cx = F(clampi)((int32_t)FLOOR(gp.x / vs), 0, (int32_t)dims[0] * 8 - 1);
    int32_t cy = F(clampi)((int32_t)FLOOR(gp.y / vs), 0, (int32_t)dims[1] * 8 - 1);
    int32_t cz = F(clampi)((int32_t)FLOOR(gp.z / vs), 0, (int32_t)dims[2] * 8 - 1);
    uint32_t flat = (uint32_t)(cx / 8) + (uint32_t)(cy / 8) * dims[0] + (uint32_t)(cz / 8) * dims[0] * dims[1];
    uint32_t slot = F(map_entry)(s, (uint64_t)o->brick_map_offset + flat);
    *slot_out = slot;
    uint32_t lv = (uint32_t)(cx % 8) + (uint32_t)(cy % 8) * 8u + (uint32_t)(cz % 8) * 64u;
    if (slot == RKF_EMPTY_SLOT || slot == RKF_INTERIOR_SLOT) return UINT64_MAX;
    *vox_out = lv;
    return (uint64_t)slot * 512u + lv;
}

static uint32_t F(voxelized_material)(const orc_ref_scene *s, V3 lp, const rkf_gpu_object *o) {
    uint32_t sl, vx;
    uint64_t i = F(nearest_voxel)(s, lp, o, &sl, &vx);
    if (i == UINT64_MAX) return 0;
    return (F(pool_word0)(s, i) >> 16) & 0x3Fu;
}

/* :182-264 */
static R F(analytical)(V3 p, const rkf_gpu_object *o) {
    R p0 = (R)o->sdf_params[0], p1 = (R)o->sdf_params[1], p2 = (R)o->sdf_params[2];
    switch (o->primitive_type) {
    case 0: return F(length)(p) - p0;
    case 1: {
        V3 q = F(mk)(FABS(p.x) - p0, FABS(p.y) - p1, FABS(p.z) - p2);
        return F(length)(F(mk)(F(maxr)(q.x, 0), F(maxr)(q.y, 0), F(maxr)(q.z, 0))) +
               F(minr)(F(maxr)(q.x, F(maxr)(q.y, q.z)), (R)0);
    }
    case 2: {
        V3 q = F(mk)(p.x, F(maxr)(FABS(p.y) - p1, (R)0), p.z);
        return F(length)(q) - p0;
    }
    case 3: {
        R qx = SQRT(p.x * p.x + p.z * p.z) - p0;
        return SQRT(qx * qx + p.y * p.y) - p1;
    }
    case 4: {
        R dx = SQRT(p.x * p.x + p.z * p.z) - p0, dy = FABS(p.y) - p1;
        R mx = F(maxr)(dx, (R)0), my = F(maxr)(dy, (R)0);
        return F(minr)(F(maxr)(dx, dy), (R)0) + SQRT(mx * mx + my * my);
    }
    case 5: {
        V3 n = F(normalize)(F(mk)(p0, p1, p2));
        return F(dot)(p, n) + (R)0;
    }
    default: return K_MAX_FLOAT;
    }
}

/* :600-634; returns distance, writes material. Indices beyond the uploaded
 * objects evaluate as SDF_TYPE_NONE. */
static R F(eval_object)(F(rctx) *c, V3 wp, uint32_t oi, uint32_t *mat) {
    const orc_ref_scene *s = c->s;
    *mat = 0;
    if (oi >= s->su.num_objects) return K_MAX_FLOAT;
    const rkf_gpu_object *o = &s->objects[oi];
    if (o->sdf_type == RKF_SDF_TYPE_NONE) return K_MAX_FLOAT;
    V3 lp = F(xform_point)(o->inverse_world, wp);
    R d;
    if (o->sdf_type == RKF_SDF_TYPE_ANALYTICAL) {
        d = F(analytical)(lp, o);
        /* evaluate_object returns the material as an f32 (:633) that the
           march converts back with u32() */
        *mat = orc_u32_via_f32(o->material_id);
    } else {
        d = F(sample_voxelized)(c, lp, o, 0);
        *mat = d < (R)o->voxel_size * (R)16.0f ? F(voxelized_material)(s, lp, o) : 0u;
    }
    R ms = F(minr)((R)o->accumulated_scale[0], F(minr)((R)o->accumulated_scale[1], (R)o->accumulated_scale[2]));
    return d * ms;
}

/* :228-235 with a software trilinear filter. */
static R F(coarse)(const orc_ref_scene *s, V3 cam_rel) {
    if (!s->coarse) return 0;
    const rkf_coarse_field_uniforms *u = &s->coarse_u;
    V3 fp = F(sub)(cam_rel, F(mk)((R)u->origin_cam_rel[0], (R)u->origin_cam_rel[1], (R)u->origin_cam_rel[2]));
    R iv = (R)u->inv_voxel_size;
    R ux = fp.x * iv / (R)u->dims[0], uy = fp.y * iv / (R)u->dims[1], uz = fp.z * iv / (R)u->dims[2];
    if (ux < (R)0 || uy < (R)0 || uz < (R)0 || ux > (R)1 || uy > (R)1 || uz > (R)1) return 0;
    R tc[3] = {ux * (R)u->dims[0] - (R)0.5f, uy * (R)u->dims[1] - (R)0.5f, uz * (R)u->dims[2] - (R)0.5f};
    int32_t i0[3], i1[3];
    R f[3];
    for (int a = 0; a < 3; a++) {
        R fl = FLOOR(tc[a]);
        f[a] = tc[a] - fl;
        int32_t ii = (int32_t)fl;
        i0[a] = F(clampi)(ii, 0, (int32_t)u->dims[a] - 1);
        i1[a] = F(clampi)(ii + 1, 0, (int32_t)u->dims[a] - 1);
    }
    uint64_t dx = u->dims[0], dxy = (uint64_t)u->dims[0] * u->dims[1];
#define CF(x, y, z) ((R)orc_f16_to_f32(s->coarse[(uint64_t)(x) + (uint64_t)(y) * dx + (uint64_t)(z) * dxy]))
    R c00 = F(mix)(CF(i0[0], i0[1], i0[2]), CF(i1[0], i0[1], i0[2]), f[0]);
    R c10 = F(mix)(CF(i0[0], i1[1], i0[2]), CF(i1[0], i1[1], i0[2]), f[0]);
    R c01 = F(mix)(CF(i0[0], i0[1], i1[2]), CF(i1[0], i0[1], i1[2]), f[0]);
    R c11 = F(mix)(CF(i0[0], i1[1], i1[2]), CF(i1[0], i1[1], i1[2]), f[0]);
#undef CF
    return F(mix)(F(mix)(c00, c10, f[1]), F(mix)(c01, c11, f[1]), f[2]);
}

typedef struct {
    R t;
    int hit;
    uint32_t mat, obj_id, obj_idx, steps;
} F(mres);

/* Shared bisection (:941-964). */
static void F(bisect)(F(rctx) *c, V3 o, V3 sd, R lo, R hi, uint32_t oi, uint32_t best_mat,
                      uint32_t best_id, F(mres) *r) {
    uint32_t bm = best_mat;
    c->flags |= ORC_FLAG_BISECT;
    for (uint32_t b = 0; b < K_BISECT; b++) {
        R mid = (lo + hi) * (R)0.5f;
        V3 mp = F(add)(o, F(muls)(sd, mid));
        uint32_t mm;
        R md = F(eval_object)(c, mp, oi, &mm);
        ORC_PATH(c->path, md < (R)0 ? 5u : 6u);
        if (md < (R)0) hi = mid;
        else { lo = mid; bm = mm; }
    }
    r->t = (lo + hi) * (R)0.5f;
    r->hit = 1;
    r->mat = bm;
    r->obj_id = best_id;
    r->obj_idx = oi;
}

/* mode: 0 tiled, 1 bvh, 2 brute. :641-985 */
static void F(march)(F(rctx) *c, V3 o, V3 dir, uint32_t px, uint32_t py, int mode, F(mres) *r) {
    const orc_ref_scene *s = c->s;
    const rkf_scene_uniforms *su = &s->su;
    r->t = K_MAX_FLOAT;
    r->hit = 0;
    r->mat = r->obj_id = r->obj_idx = 0;
    r->steps = 0;
    uint32_t count = 0, base = 0;
    if (mode == 0) {
        uint32_t ntx = (s->width + 15u) / 16u;
        uint32_t tid = (py / 16u) * ntx + px / 16u;
        count = tid < s->tiles_x * s->tiles_y ? s->tile_counts[tid] : 0u;
        if (count > RKF_TILE_MAX_OBJECTS) count = RKF_TILE_MAX_OBJECTS;
        if (count == 0) return;
        base = tid * RKF_TILE_MAX_OBJECTS;
    } else if (mode == 1) {
        if (su->num_objects == 0) return;
    }
    V3 sd = F(mk)(FABS(dir.x) < (R)1e-10f ? (R)1e-10f : dir.x, FABS(dir.y) < (R)1e-10f ? (R)1e-10f : dir.y,
                  FABS(dir.z) < (R)1e-10f ? (R)1e-10f : dir.z);
    V3 cam = F(mk)((R)s->cam.position[0], (R)s->cam.position[1], (R)s->cam.position[2]);
    R t = 0, prev_t = 0, prev_dist = K_MAX_FLOAT;
    uint32_t near = 0;
    for (uint32_t step = 0; step < K_HARD_MAX; step++) {
        if (t > (R)su->max_distance || near >= su->max_steps) break;
        r->steps++;
        V3 pos = F(add)(o, F(muls)(sd, t));
        if (mode == 0) {
            R cd = F(coarse)(s, F(sub)(pos, cam));
            if (cd > K_COARSE_NEAR) {
                ORC_PATH(c->path, 1u);
                prev_dist = K_MAX_FLOAT;
                t += cd;
                continue;
            }
        }
        R min_dist = K_MAX_FLOAT;
        uint32_t best_mat = 0, best_id = 0, best_idx = 0;
        if (mode == 0) {
            for (uint32_t i = 0; i < count; i++) {
                uint32_t oi = s->tile_idx[base + i];
                uint32_t m;
                R d = F(eval_object)(c, pos, oi, &m);
                if (d < min_dist) {
                    min_dist = d;
                    best_mat = m;
                    best_id = oi < su->num_objects ? s->objects[oi].object_id : 0u;
                    best_idx = oi;
                }
            }
        } else if (mode == 1) {
            uint32_t stack[32];
            uint32_t sp = 1;
            stack[0] = 0;
            while (sp > 0) {
                uint32_t ni = stack[--sp];
                if (ni >= s->n_bvh) continue;
                const rkf_bvh_node *n = &s->bvh[ni];
                V3 nmin = F(mk)((R)n->aabb_min[0], (R)n->aabb_min[1], (R)n->aabb_min[2]);
                V3 nmax = F(mk)((R)n->aabb_max[0], (R)n->aabb_max[1], (R)n->aabb_max[2]);
                V3 cl = F(mk)(F(clampr)(pos.x, nmin.x, nmax.x), F(clampr)(pos.y, nmin.y, nmax.y),
                              F(clampr)(pos.z, nmin.z, nmax.z));
                R bd = F(length)(F(sub)(cl, pos));
                if (bd > min_dist) continue;
                if (n->left == RKF_BVH_INVALID) {
                    uint32_t oi = n->right_or_object;
                    if (oi < su->num_objects) {
                        uint32_t m;
                        R d = F(eval_object)(c, pos, oi, &m);
                        if (d < min_dist) {
                            min_dist = d;
                            best_mat = m;
                            best_id = s->objects[oi].object_id;
                            best_idx = oi;
                        }
                    }
                } else if (sp < 31u) {
                    stack[sp++] = n->left;
                    stack[sp++] = n->right_or_object;
                }
            }
        } else {
            for (uint32_t i = 0; i < su->num_objects; i++) {
                uint32_t m;
                R d = F(eval_object)(c, pos, i, &m);
                if (d < min_dist) {
                    min_dist = d;
                    best_mat = m;
                    best_id = s->objects[i].object_id;
                    best_idx = i;
                }
            }
        }
        if (min_dist < (R)0 && prev_dist > F(hit_eps)(&s->ext, su, prev_t)) {
            F(bisect)(c, o, sd, prev_t, t, best_idx, best_mat, best_id, r);
            return;
        }
        if (min_dist < F(hit_eps)(&s->ext, su, t)) {
            r->t = t;
            r->hit = 1;
            r->mat = best_mat;
            r->obj_id = best_id;
            r->obj_idx = best_idx;
            return;
        }
        ORC_PATH(c->path, (min_dist < K_NEAR ? 3u : 2u) + 8u * best_idx);
        if (min_dist < K_NEAR) near++;
        prev_t = t;
        prev_dist = min_dist;
        t += F(maxr)(min_dist, K_MIN_STEP);
    }
}

static R F(gradient_sample)(F(rctx) *c, V3 lp, const rkf_gpu_object *o) {
    return F(sample_voxelized)(c, lp, o, 1);
}

/* :1003-1032 */
static V3 F(normal)(F(rctx) *c, V3 pos, uint32_t oi) {
    const orc_ref_scene *s = c->s;
    const rkf_gpu_object *o = &s->objects[oi];
    if (o->sdf_type == RKF_SDF_TYPE_VOXELIZED) {
        V3 lp = F(xform_point)(o->inverse_world, pos);
        R e = (R)o->voxel_size * (R)2.0f;
        R gx = F(gradient_sample)(c, F(add)(lp, F(mk)(e, 0, 0)), o) - F(gradient_sample)(c, F(sub)(lp, F(mk)(e, 0, 0)), o);
        R gy = F(gradient_sample)(c, F(add)(lp, F(mk)(0, e, 0)), o) - F(gradient_sample)(c, F(sub)(lp, F(mk)(0, e, 0)), o);
        R gz = F(gradient_sample)(c, F(add)(lp, F(mk)(0, 0, e)), o) - F(gradient_sample)(c, F(sub)(lp, F(mk)(0, 0, e)), o);
        V3 wg = F(xform_grad)(o->inverse_world, F(mk)(gx, gy, gz));
        R len = F(length)(wg);
        if (len < (R)1e-10f) return F(mk)(0, 1, 0);
        return F(mk)(wg.x / len, wg.y / len, wg.z / len);
    }
    R e = (R)s->su.hit_threshold * (R)10.0f;
    uint32_t m;
    R nx = F(eval_object)(c, F(add)(pos, F(mk)(e, 0, 0)), oi, &m) - F(eval_object)(c, F(sub)(pos, F(mk)(e, 0, 0)), oi, &m);
    R ny = F(eval_object)(c, F(add)(pos, F(mk)(0, e, 0)), oi, &m) - F(eval_object)(c, F(sub)(pos, F(mk)(0, e, 0)), oi, &m);
    R nz = F(eval_object)(c, F(add)(pos, F(mk)(0, 0, e)), oi, &m) - F(eval_object)(c, F(sub)(pos, F(mk)(0, 0, e)), oi, &m);
    return F(normalize)(F(mk)(nx, ny, nz));
}

/* Ray generation :1046-1053. */
static void F(camera_ray)(const rkf_camera_uniforms *cam, uint32_t w, uint32_t h, uint32_t px, uint32_t py,
                          V3 *o, V3 *dir) {
    R ux = ((R)px + (R)0.5f + (R)cam->jitter[0]) / (R)w;
    R uy = ((R)py + (R)0.5f + (R)cam->jitter[1]) / (R)h;
    R nx = ux * (R)2.0f - (R)1.0f, ny = (R)1.0f - uy * (R)2.0f;
    *o = F(mk)((R)cam->position[0], (R)cam->position[1], (R)cam->position[2]);
    V3 f = F(mk)((R)cam->forward[0], (R)cam->forward[1], (R)cam->forward[2]);
    V3 rt = F(mk)((R)cam->right[0], (R)cam->right[1], (R)cam->right[2]);
    V3 up = F(mk)((R)cam->up[0], (R)cam->up[1], (R)cam->up[2]);
    *dir = F(normalize)(F(add)(F(add)(f, F(muls)(rt, nx)), F(muls)(up, ny)));
}

/* Motion vector :1096-1103. */
static void F(motion)(const rkf_camera_uniforms *cam, uint32_t w, uint32_t h, uint32_t px, uint32_t py,
                      V3 hp, R *mx, R *my) {
    const float *m = cam->prev_vp;
    R cx = (R)m[0] * hp.x + (R)m[4] * hp.y + (R)m[8] * hp.z + (R)m[12];
    R cy = (R)m[1] * hp.x + (R)m[5] * hp.y + (R)m[9] * hp.z + (R)m[13];
    R cw = (R)m[3] * hp.x + (R)m[7] * hp.y + (R)m[11] * hp.z + (R)m[15];
    *mx = 0;
    *my = 0;
    if (cw > (R)0) {
        R nx = cx / cw, ny = cy / cw;
        R pu = nx * (R)0.5f + (R)0.5f, pv = (R)0.5f - ny * (R)0.5f;
        R cu = ((R)px + (R)0.5f) / (R)w, cv = ((R)py + (R)0.5f) / (R)h;
        *mx = cu - pu;
        *my = cv - pv;
    }
}

static void F(ref_pixel)(const orc_ref_scene *s, uint64_t p, orc_ray_result *out) {
    F(rctx) c = {s, 0, 0, 0};
    uint32_t px = (uint32_t)(p % s->width), py = (uint32_t)(p / s->width);
    V3 o, dir;
    F(camera_ray)(&s->cam, s->width, s->height, px, py, &o, &dir);
    int mode;
    uint32_t st = s->ext.strategy;
    if (st == RKFM_STRATEGY_TILED) mode = 0;
    else if (st == RKFM_STRATEGY_BVH) mode = 1;
    else if (st == RKFM_STRATEGY_BRUTE) mode = 2;
    else mode = s->tile_counts ? 0 : (s->n_bvh ? 1 : 2);
    F(mres) r;
    F(march)(&c, o, dir, px, py, mode, &r);
    memset(out, 0, sizeof(*out));
    out->steps = r.steps;
    out->brick_slot = RKF_EMPTY_SLOT;
    out->voxel_index = 0xFFFFFFFFu;
    out->cell_index = 0xFFFFFFFFu;
    if (!r.hit) {
        out->evals = c.evals;
        out->flags = c.flags;
        out->path = c.path;
        out->t = (double)K_MAX_FLOAT;
        return;
    }
    V3 hp = F(add)(o, F(muls)(dir, r.t));
    uint32_t evals_march = c.evals;
    V3 n = F(normal)(&c, hp, r.obj_idx);
    const rkf_gpu_object *ho = &s->objects[r.obj_idx];
    uint32_t sec = 0, blend = 0;
    R sdf_at = 0;
    out->cell_index = r.obj_idx;
    if (ho->sdf_type == RKF_SDF_TYPE_VOXELIZED) {
        V3 lh = F(xform_point)(ho->inverse_world, hp);
        uint32_t sl, vx;
        uint64_t i = F(nearest_voxel)(s, lh, ho, &sl, &vx);
        out->brick_slot = sl;
        out->voxel_index = vx;
        if (i != UINT64_MAX) {
            sec = (F(pool_word0)(s, i) >> 22) & 0x3Fu;
            blend = (F(pool_word1)(s, i) >> 24) & 0xFFu;
        }
        sdf_at = F(sample_voxelized)(&c, lh, ho, 0);
    }
    out->evals = evals_march;
    out->path = c.path;
    out->flags = c.flags | (ho->sdf_type == RKF_SDF_TYPE_ANALYTICAL ? ORC_FLAG_ANALYTIC_HIT : 0u);
    out->hit = 1;
    out->t = (double)r.t;
    out->pos[0] = (double)hp.x;
    out->pos[1] = (double)hp.y;
    out->pos[2] = (double)hp.z;
    out->normal[0] = (double)n.x;
    out->normal[1] = (double)n.y;
    out->normal[2] = (double)n.z;
    out->material_id = r.mat;
    out->obj_idx = r.obj_idx;
    out->object_id = r.obj_id;
    out->packed_material = (r.mat & 0x3Fu) | (sec << 6) | (blend << 12) | (r.obj_id << 24);
    R mx, my;
    F(motion)(&s->cam, s->width, s->height, px, py, hp, &mx, &my);
    out->motion[0] = (double)mx;
    out->motion[1] = (double)my;
    out->sdf_at_hit = (double)sdf_at;
}

/* ================================================================== */
/* Sparse two-level scenes                                              */
/* ================================================================== */
/*
 * Field: identical to sample_voxelized for one object whose dense brick map is
 * the expansion of the index (grid_pos = p - origin, no geometry early-out),
 * except that grid coordinates multiply by 1/vs (rounded once to the working
 * precision) where the WGSL divides by vs: WGSL's f32 division is only
 * specified to 2.5 ULP, and a multiply costs one instruction on the GPU
 * instead of a correctly rounded division's dozen.
 * L2 cell = 8^3 bricks; l2[cell] = L1 block or EMPTY; l1[block*512 + brick in
 * cell] = slot / EMPTY / INTERIOR; dist[slot*512 + voxel] binary16.
 *
 * Free-space DDA: let g = (p - origin)*(1/vs) - 0.5 (the trilinear coordinate) and
 * b0 = floor(floor(g)/8). A sample's 8 corners lie in bricks {b0, b0+1}^3, so
 * if all of those are EMPTY (b0 "free") the field is the EMPTY constant
 * 8 vs over the whole region floor(g/8) == b0; if every brick of an L2 cell is
 * free, over floor(g/64) == cell, and the cell's octant extent E >= 1 for the
 * ray's octant (oct[cell], see orc_sparse_oct) extends that region to the
 * cube of cells q + [0, E) taken in the ray's direction on each axis; in a
 * cell that is not entirely free, a free brick's octant extent E_b (bext,
 * see orc_sparse_bext) extends it to the cube of bricks b0 + [0, E_b). Sphere
 * tracing through such
 * a region only takes positive, non-near steps; the march instead advances to
 * the region's exit t_x along the ray, sets prev_t = t_x, prev_dist = 8 vs
 * (the field value there) and resumes at t_x + vs/256. Each skip counts as one
 * loop iteration. It applies only where 0 <= floor(g) <= total-2 per axis and
 * the position is inside the clamp range of the grid.
 */
typedef struct {
    const orc_sparse_scene *s;
    uint32_t evals;
    uint32_t flags; /* ORC_FLAG_* */
    uint32_t nb[3];
    int32_t tot[3];
} F(sctx);

static inline uint32_t F(sp_slot)(const F(sctx) *c, int32_t cx, int32_t cy, int32_t cz) {
    const orc_sparse_scene *s = c->s;
    uint32_t bx = (uint32_t)cx >> 3, by = (uint32_t)cy >> 3, bz = (uint32_t)cz >> 3;
    uint32_t qx = bx >> 3, qy = by >> 3, qz = bz >> 3;
    uint64_t cell = (uint64_t)qx + (uint64_t)qy * s->cells[0] + (uint64_t)qz * s->cells[0] * s->cells[1];
    uint32_t blk = s->l2[cell];
    if (blk == RKF_EMPTY_SLOT) return RKF_EMPTY_SLOT;
    return s->l1[(uint64_t)blk * 512u + (bx & 7u) + (by & 7u) * 8u + (bz & 7u) * 64u];
}

static R F(sp_voxel)(F(sctx) *c, int32_t cx, int32_t cy, int32_t cz, R vs, int grad) {
    cx = F(clampi)(cx, 0, c->tot[0] - 1);
    cy = F(clampi)(cy, 0, c->tot[1] - 1);
    cz = F(clampi)(cz, 0, c->tot[2] - 1);
    uint32_t slot = F(sp_slot)(c, cx, cy, cz);
    if (slot == RKF_EMPTY_SLOT) return vs * (R)8.0f;
    if (slot == RKF_INTERIOR_SLOT) {
        c->flags |= grad ? ORC_FLAG_INTERIOR_GRAD : ORC_FLAG_INTERIOR_STEP;
        return grad ? -(vs * (R)8.0f) : -(vs * (R)2.0f);
    }
    uint64_t idx = (uint64_t)slot * 512u + (uint32_t)(cx & 7) + (uint32_t)(cy & 7) * 8u + (uint32_t)(cz & 7) * 64u;
    return (R)orc_f16_to_f32(c->s->dist[idx]);
}

static R F(sp_sample)(F(sctx) *c, V3 p, int grad, int count) {
    const orc_sparse_scene *s = c->s;
    R vs = (R)s->voxel_size;
    R be = vs * (R)8.0f;
    V3 gs = F(mk)((R)c->nb[0] * be, (R)c->nb[1] * be, (R)c->nb[2] * be);
    V3 gp = F(sub)(p, F(mk)((R)s->origin[0], (R)s->origin[1], (R)s->origin[2]));
    R lo = vs * (R)0.01f;
    V3 cl = F(mk)(F(clampr)(gp.x, lo, gs.x - lo), F(clampr)(gp.y, lo, gs.y - lo), F(clampr)(gp.z, lo, gs.z - lo));
    R outside = F(length)(F(sub)(gp, cl));
    if (outside > be * (R)2.0f) return outside;
    R ivs = (R)1.0f / vs;
    V3 vc = F(mk)(cl.x * ivs - (R)0.5f, cl.y * ivs - (R)0.5f, cl.z * ivs - (R)0.5f);
    int32_t x0 = (int32_t)FLOOR(vc.x), y0 = (int32_t)FLOOR(vc.y), z0 = (int32_t)FLOOR(vc.z);
    R tx = vc.x - (R)x0, ty = vc.y - (R)y0, tz = vc.z - (R)z0;
    if (count) c->evals++;
    R c000 = F(sp_voxel)(c, x0, y0, z0, vs, grad), c100 = F(sp_voxel)(c, x0 + 1, y0, z0, vs, grad);
    R c010 = F(sp_voxel)(c, x0, y0 + 1, z0, vs, grad), c110 = F(sp_voxel)(c, x0 + 1, y0 + 1, z0, vs, grad);
    R c001 = F(sp_voxel)(c, x0, y0, z0 + 1, vs, grad), c101 = F(sp_voxel)(c, x0 + 1, y0, z0 + 1, vs, grad);
    R c011 = F(sp_voxel)(c, x0, y0 + 1, z0 + 1, vs, grad), c111 = F(sp_voxel)(c, x0 + 1, y0 + 1, z0 + 1, vs, grad);
    R c00 = F(mix)(c000, c100, tx), c10 = F(mix)(c010, c110, tx);
    R c01 = F(mix)(c001, c101, tx), c11 = F(mix)(c011, c111, tx);
    R c0 = F(mix)(c00, c10, ty), c1 = F(mix)(c01, c11, ty);
    return F(mix)(c0, c1, tz) + outside;
}

/* Nearest voxel (floor(grid_pos / vs)), as :540-596. Writes slot/voxel/cell. */
static int64_t F(sp_nearest)(const F(sctx) *c, V3 p, uint32_t *slot_o, uint32_t *vox_o, uint32_t *cell_o) {
    const orc_sparse_scene *s = c->s;
    R vs = (R)s->voxel_size;
    R be = vs * (R)8.0f;
    V3 gs = F(mk)((R)c->nb[0] * be, (R)c->nb[1] * be, (R)c->nb[2] * be);
    V3 gp = F(sub)(p, F(mk)((R)s->origin[0], (R)s->origin[1], (R)s->origin[2]));
    *slot_o = RKF_EMPTY_SLOT;
    *vox_o = 0xFFFFFFFFu;
    *cell_o = 0xFFFFFFFFu;
    if (gp.x < (R)0 || gp.y < (R)0 || gp.z < (R)0 || gp.x >= gs.x || gp.y >= gs.y || gp.z >= gs.z) return -1;
    R ivs = (R)1.0f / vs;
    int32_t cx = F(clampi)((int32_t)FLOOR(gp.x * ivs), 0, c->tot[0] - 1);
    int32_t cy = F(clampi)((int32_t)FLOOR(gp.y * ivs), 0, c->tot[1] - 1);
    int32_t cz = F(clampi)((int32_t)FLOOR(gp.z * ivs), 0, c->tot[2] - 1);
    uint32_t slot = F(sp_slot)(c, cx, cy, cz);
    *cell_o = ((uint32_t)cx >> 6) + ((uint32_t)cy >> 6) * s->cells[0] + ((uint32_t)cz >> 6) * s->cells[0] * s->cells[1];
    *slot_o = slot;
    uint32_t lv = (uint32_t)(cx & 7) + (uint32_t)(cy & 7) * 8u + (uint32_t)(cz & 7) * 64u;
    if (slot == RKF_EMPTY_SLOT || slot == RKF_INTERIOR_SLOT) return -1;
    *vox_o = lv;
    return (int64_t)((uint64_t)slot * 512u + lv);
}

static uint32_t F(sp_material)(const F(sctx) *c, V3 p) {
    uint32_t a, b, d;
    int64_t i = F(sp_nearest)(c, p, &a, &b, &d);
    return i < 0 ? 0u : (uint32_t)(c->s->mat[i] & 0x3Fu);
}

/* DDA: returns 1 and the exit parameter when the sample at `pos` lies in a
 * free region. */
static int F(sp_free)(const F(sctx) *c, V3 pos, V3 sd, V3 inv, R t, R *t_exit) {
    const orc_sparse_scene *s = c->s;
    R vs = (R)s->voxel_size;
    R be = vs * (R)8.0f;
    V3 gs = F(mk)((R)c->nb[0] * be, (R)c->nb[1] * be, (R)c->nb[2] * be);
    V3 gp = F(sub)(pos, F(mk)((R)s->origin[0], (R)s->origin[1], (R)s->origin[2]));
    R lo = vs * (R)0.01f;
    if (gp.x < lo || gp.y < lo || gp.z < lo || gp.x > gs.x - lo || gp.y > gs.y - lo || gp.z > gs.z - lo) return 0;
    R ivs = (R)1.0f / vs;
    R g[3] = {gp.x * ivs - (R)0.5f, gp.y * ivs - (R)0.5f, gp.z * ivs - (R)0.5f};
    int32_t v0[3];
    for (int a = 0; a < 3; a++) {
        v0[a] = (int32_t)FLOOR(g[a]);
        if (v0[a] < 0 || v0[a] > c->tot[a] - 2) return 0;
    }
    uint32_t b[3] = {(uint32_t)v0[0] >> 3, (uint32_t)v0[1] >> 3, (uint32_t)v0[2] >> 3};
    uint32_t q[3] = {b[0] >> 3, b[1] >> 3, b[2] >> 3};
    uint64_t cell = (uint64_t)q[0] + (uint64_t)q[1] * s->cells[0] + (uint64_t)q[2] * s->cells[0] * s->cells[1];
    R lo_b[3], hi_b[3];
    uint32_t octant = (sd.x < (R)0 ? 1u : 0u) | (sd.y < (R)0 ? 2u : 0u) | (sd.z < (R)0 ? 4u : 0u);
    uint32_t e = (s->oct[cell] >> (4u * octant)) & 15u;
    if (e > 0) {
        /* the cube of cells q + [0, e) in the travel direction of each axis is free */
        for (int a = 0; a < 3; a++) {
            lo_b[a] = ((R)q[a] - (R)(e - 1u)) * (R)64.0f;
            hi_b[a] = ((R)q[a] + (R)e) * (R)64.0f;
        }
    } else {
        uint32_t lb = (b[0] & 7u) + (b[1] & 7u) * 8u + (b[2] & 7u) * 64u;
        uint32_t eb = (s->bext[cell * 512u + lb] >> (4u * octant)) & 15u;
        if (eb == 0) return 0;
        /* the cube of bricks b + [0, eb) in the travel direction of each axis is free */
        for (int a = 0; a < 3; a++) {
            lo_b[a] = ((R)b[a] - (R)(eb - 1u)) * (R)8.0f;
            hi_b[a] = ((R)b[a] + (R)eb) * (R)8.0f;
        }
    }
    R dv[3] = {sd.x, sd.y, sd.z};
    R iv[3] = {inv.x, inv.y, inv.z};
    R dt = K_MAX_FLOAT;
    for (int a = 0; a < 3; a++) {
        R bound = dv[a] > (R)0 ? hi_b[a] : lo_b[a];
        R da = (bound - g[a]) * iv[a];
        dt = F(minr)(dt, da);
    }
    *t_exit = t + F(maxr)(dt, (R)0);
    return 1;
}

typedef struct {
    R t, mat_d;
    V3 mat_p;
    int hit;
    uint32_t steps;
    uint32_t path; /* ORC_PATH hash of the march's discrete decisions */
} F(sres);

/* Sparse march: ray_march_tiled's step rules (:905-982) over one field. */
/* Tile intervals (RKFM_MARCH_TILE_CULL; no reference counterpart). For
 * every L2 cell holding an L1 block: its box widened by 9 vs (25 vs on sides
 * on the grid boundary), tn / tf = distance from the camera to the box / to
 * its farthest corner, and the rectangle of 8x8 tiles whose pixel rays can
 * meet it (corner projections through the inverted camera basis, 2 pixels
 * wider; boxes behind the camera or outside one frustum side plane are
 * dropped, boxes straddling the camera plane cover every tile). Per tile:
 * lo = min tn, hi = max tf (+inf / 0 when no box). */
static int F(tile_cell)(const orc_sparse_scene *s, const V3 *tc, R w, R h, R jx, R jy, uint32_t cell, R maxd,
                        uint32_t tx, uint32_t ty, uint32_t *rect, R *tn, R *tf) {
    /* tc: o, ia, ib, ic */
    uint32_t q[3] = {cell % s->cells[0], (cell / s->cells[0]) % s->cells[1], cell / (s->cells[0] * s->cells[1])};
    R vs = (R)s->voxel_size;
    R cw = vs * (R)64.0f, mg = vs * (R)9.0f, mb = vs * (R)9.0f + vs * (R)16.0f;
    R o[3] = {tc[0].x, tc[0].y, tc[0].z};
    R lo[3], hi[3], dn2 = 0, df2 = 0;
    for (int a = 0; a < 3; a++) {
        lo[a] = (R)s->origin[a] + (R)q[a] * cw - (q[a] == 0u ? mb : mg);
        hi[a] = (R)s->origin[a] + (R)(q[a] + 1u) * cw + (q[a] + 1u == s->cells[a] ? mb : mg);
        R dl = lo[a] - o[a], dh = o[a] - hi[a];
        R n = FMAX(FMAX(dl, dh), (R)0);
        R fa = FMAX(FABS(dl), FABS(dh));
        dn2 = dn2 + n * n;
        df2 = df2 + fa * fa;
    }
    *tn = SQRT(dn2);
    *tf = SQRT(df2);
    if (*tn > maxd) return 0;
    int all_front = 1, any_front = 0;
    uint32_t outside = 15u;
    R nx0 = (R)INFINITY, nx1 = -(R)INFINITY, ny0 = (R)INFINITY, ny1 = -(R)INFINITY;
    for (uint32_t k = 0; k < 8u; k++) {
        V3 v = F(sub)(F(mk)((k & 1u) ? hi[0] : lo[0], (k & 2u) ? hi[1] : lo[1], (k & 4u) ? hi[2] : lo[2]), tc[0]);
        R a = F(dot)(tc[1], v), b = F(dot)(tc[2], v), cc = F(dot)(tc[3], v);
        R sl = a + FABS(a) * (R)1e-3f;
        if (!(b > sl)) outside &= ~1u;
        if (!(-b > sl)) outside &= ~2u;
        if (!(cc > sl)) outside &= ~4u;
        if (!(-cc > sl)) outside &= ~8u;
        if (a > (R)0) {
            any_front = 1;
            R nx = b / a, ny = cc / a;
            nx0 = FMIN(nx0, nx);
            nx1 = FMAX(nx1, nx);
            ny0 = FMIN(ny0, ny);
            ny1 = FMAX(ny1, ny);
        } else {
            all_front = 0;
        }
    }
    if (outside) return 0;
    if (!any_front && *tn > (R)0) return 0;
    if (!all_front) {
        rect[0] = 0; rect[1] = tx - 1u; rect[2] = 0; rect[3] = ty - 1u;
        return 1;
    }
    R fx0 = (nx0 + (R)1.0f) * (R)0.5f * w - (R)0.5f - jx - (R)2.0f;
    R fx1 = (nx1 + (R)1.0f) * (R)0.5f * w - (R)0.5f - jx + (R)2.0f;
    R fy0 = ((R)1.0f - ny1) * (R)0.5f * h - (R)0.5f - jy - (R)2.0f;
    R fy1 = ((R)1.0f - ny0) * (R)0.5f * h - (R)0.5f - jy + (R)2.0f;
    fx0 = FMIN(FMAX(fx0, (R)-16.0f), w + (R)16.0f);
    fx1 = FMIN(FMAX(fx1, (R)-16.0f), w + (R)16.0f);
    fy0 = FMIN(FMAX(fy0, (R)-16.0f), h + (R)16.0f);
    fy1 = FMIN(FMAX(fy1, (R)-16.0f), h + (R)16.0f);
    int32_t ix0 = (int32_t)FLOOR(fx0 * (R)0.125f), ix1 = (int32_t)FLOOR(fx1 * (R)0.125f);
    int32_t iy0 = (int32_t)FLOOR(fy0 * (R)0.125f), iy1 = (int32_t)FLOOR(fy1 * (R)0.125f);
    if (ix1 < 0 || iy1 < 0 || ix0 > (int32_t)tx - 1 || iy0 > (int32_t)ty - 1) return 0;
    rect[0] = (uint32_t)(ix0 < 0 ? 0 : ix0);
    rect[1] = (uint32_t)(ix1 > (int32_t)tx - 1 ? (int32_t)tx - 1 : ix1);
    rect[2] = (uint32_t)(iy0 < 0 ? 0 : iy0);
    rect[3] = (uint32_t)(iy1 > (int32_t)ty - 1 ? (int32_t)ty - 1 : iy1);
    return 1;
}

static V3 F(cross3)(V3 a, V3 b) {
    return F(mk)(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

/* iv: 2 values (lo, hi) per tile, row-major over ceil(w/8) x ceil(h/8);
 * cnt (or NULL): per tile, the number of cells binned into it (the kernel's
 * launch-order key). */
static void F(sp_tiles)(const orc_sparse_scene *s, const rkf_camera_uniforms *cam, R maxd, uint32_t w, uint32_t h,
                        R *iv, uint32_t *cnt) {
    uint32_t tx = (w + 7u) / 8u, ty = (h + 7u) / 8u;
    for (uint64_t i = 0; i < (uint64_t)tx * ty; i++) {
        iv[2 * i] = (R)INFINITY;
        iv[2 * i + 1] = 0;
        if (cnt) cnt[i] = 0;
    }
    V3 f = F(mk)((R)cam->forward[0], (R)cam->forward[1], (R)cam->forward[2]);
    V3 r = F(mk)((R)cam->right[0], (R)cam->right[1], (R)cam->right[2]);
    V3 u = F(mk)((R)cam->up[0], (R)cam->up[1], (R)cam->up[2]);
    V3 ru = F(cross3)(r, u), uf = F(cross3)(u, f), fr = F(cross3)(f, r);
    R det = F(dot)(f, ru);
    V3 tc[4];
    tc[0] = F(mk)((R)cam->position[0], (R)cam->position[1], (R)cam->position[2]);
    tc[1] = F(mk)(ru.x / det, ru.y / det, ru.z / det);
    tc[2] = F(mk)(uf.x / det, uf.y / det, uf.z / det);
    tc[3] = F(mk)(fr.x / det, fr.y / det, fr.z / det);
    uint64_t ncell = (uint64_t)s->cells[0] * s->cells[1] * s->cells[2];
    for (uint64_t cell = 0; cell < ncell; cell++) {
        if (s->l2[cell] == RKF_EMPTY_SLOT) continue;
        uint32_t rect[4];
        R tn, tf;
        if (!F(tile_cell)(s, tc, (R)w, (R)h, (R)cam->jitter[0], (R)cam->jitter[1], (uint32_t)cell, maxd, tx, ty, rect,
                          &tn, &tf))
            continue;
        for (uint32_t y = rect[2]; y <= rect[3]; y++)
            for (uint32_t x = rect[0]; x <= rect[1]; x++) {
                R *t = iv + 2 * ((uint64_t)y * tx + x);
                t[0] = FMIN(t[0], tn);
                t[1] = FMAX(t[1], tf);
                if (cnt) cnt[(uint64_t)y * tx + x]++;
            }
    }
}

/* iv: the ray's tile (lo, hi) or NULL. */
static void F(sp_march)(F(sctx) *c, V3 o, V3 dir, const rkf_scene_uniforms *su, const rkfm_scene_ext *ext,
                        const R *iv, F(sres) *r) {
    R vs = (R)c->s->voxel_size;
    r->hit = 0;
    r->t = K_MAX_FLOAT;
    r->steps = 0;
    r->path = 0;
    V3 sd = F(mk)(FABS(dir.x) < (R)1e-10f ? (R)1e-10f : dir.x, FABS(dir.y) < (R)1e-10f ? (R)1e-10f : dir.y,
                  FABS(dir.z) < (R)1e-10f ? (R)1e-10f : dir.z);
    R t = 0, prev_t = 0, prev_dist = K_MAX_FLOAT;
    R nudge = vs * (R)0.00390625f;
    V3 inv = F(mk)(vs / sd.x, vs / sd.y, vs / sd.z);
    uint32_t near = 0;
    R tmax = (R)su->max_distance;
    if (iv) {
        /* outside [lo, hi] the field along the ray is the free value: one
           skip to lo, a miss past hi */
        tmax = F(minr)(tmax, iv[1]);
        if (iv[0] > t) {
            r->steps = 1;
            prev_t = iv[0];
            prev_dist = vs * (R)8.0f;
            t = iv[0] + nudge;
        }
    }
    for (;;) {
        if (r->steps >= K_HARD_MAX || t > tmax || near >= su->max_steps) break;
        r->steps++;
        V3 pos = F(add)(o, F(muls)(sd, t));
        R tx;
        if (F(sp_free)(c, pos, sd, inv, t, &tx)) {
            ORC_PATH(r->path, 1u);
            prev_t = tx;
            prev_dist = vs * (R)8.0f;
            t = tx + nudge;
            continue;
        }
        R d = F(sp_sample)(c, pos, 0, 1);
        if (d < (R)0 && prev_dist > F(hit_eps)(ext, su, prev_t)) {
            R lo = prev_t, hi = t, mt = t, md = d;
            c->flags |= ORC_FLAG_BISECT;
            for (uint32_t b = 0; b < K_BISECT; b++) {
                R mid = (lo + hi) * (R)0.5f;
                R mdd = F(sp_sample)(c, F(add)(o, F(muls)(sd, mid)), 0, 1);
                ORC_PATH(r->path, mdd < (R)0 ? 5u : 6u);
                if (mdd < (R)0) hi = mid;
                else { lo = mid; mt = mid; md = mdd; }
            }
            r->t = (lo + hi) * (R)0.5f;
            r->hit = 1;
            r->mat_p = F(add)(o, F(muls)(sd, mt));
            r->mat_d = md;
            return;
        }
        if (d < F(hit_eps)(ext, su, t)) {
            r->t = t;
            r->hit = 1;
            r->mat_p = pos;
            r->mat_d = d;
            return;
        }
        ORC_PATH(r->path, d < K_NEAR ? 3u : 2u);
        if (d < K_NEAR) near++;
        prev_t = t;
        prev_dist = d;
        t += F(maxr)(d, K_MIN_STEP);
    }
}

/* soft_shadow (shade_common_shading.wgsl:9-33) over the sparse field, with
 * the same free-space DDA (a skip counts as one iteration). */
static R F(sp_shadow)(F(sctx) *c, V3 o, V3 L, R maxd, R k, uint32_t *steps) {
    R vs = (R)c->s->voxel_size;
    R nudge = vs * (R)0.00390625f;
    R shadow = 1, t = (R)0.005f;
    V3 inv = F(mk)(vs / L.x, vs / L.y, vs / L.z);
    uint32_t near = 0;
    *steps = 0;
    for (uint32_t i = 0; i < 256u; i++) {
        if (near >= 32u || t > maxd) break;
        (*steps)++;
        V3 p = F(add)(o, F(muls)(L, t));
        R tx;
        if (F(sp_free)(c, p, L, inv, t, &tx)) { t = tx + nudge; continue; }
        R d = F(sp_sample)(c, p, 0, 1);
        if (d < (R)0.005f) return 0;
        if (d < (R)0.02f) {
            shadow = F(minr)(shadow, k * d / t);
            near++;
        }
        t += F(maxr)(d, (R)0.005f);
    }
    return F(clampr)(shadow, 0, 1);
}

static void F(sp_init)(F(sctx) *c, const orc_sparse_scene *s) {
    c->s = s;
    c->evals = 0;
    c->flags = 0;
    for (int a = 0; a < 3; a++) {
        c->nb[a] = s->cells[a] * 8u;
        c->tot[a] = (int32_t)(s->cells[a] * 64u);
    }
}

static void F(sp_pixel)(const orc_sparse_scene *s, const rkf_camera_uniforms *cam, const rkf_scene_uniforms *su,
                        const rkfm_scene_ext *ext, uint32_t w, uint32_t h, const R *tiles, uint64_t p,
                        orc_ray_result *out) {
    F(sctx) c;
    F(sp_init)(&c, s);
    uint32_t px = (uint32_t)(p % w), py = (uint32_t)(p / w);
    V3 o, dir;
    F(camera_ray)(cam, w, h, px, py, &o, &dir);
    F(sres) r;
    const R *iv = tiles ? tiles + 2 * ((uint64_t)(py >> 3) * ((w + 7u) / 8u) + (px >> 3)) : NULL;
    F(sp_march)(&c, o, dir, su, ext, iv, &r);
    memset(out, 0, sizeof(*out));
    out->steps = r.steps;
    out->brick_slot = RKF_EMPTY_SLOT;
    out->voxel_index = 0xFFFFFFFFu;
    out->cell_index = 0xFFFFFFFFu;
    out->evals = c.evals;
    out->flags = c.flags;
    out->path = r.path;
    if (!r.hit) {
        out->t = (double)K_MAX_FLOAT;
        out->shadow = 0;
        return;
    }
    V3 hp = F(add)(o, F(muls)(dir, r.t));
    uint32_t mat = 0;
    if (r.mat_d < (R)s->voxel_size * (R)16.0f) mat = F(sp_material)(&c, r.mat_p);
    R e = (R)s->voxel_size * (R)2.0f;
    R gx = F(sp_sample)(&c, F(add)(hp, F(mk)(e, 0, 0)), 1, 0) - F(sp_sample)(&c, F(sub)(hp, F(mk)(e, 0, 0)), 1, 0);
    R gy = F(sp_sample)(&c, F(add)(hp, F(mk)(0, e, 0)), 1, 0) - F(sp_sample)(&c, F(sub)(hp, F(mk)(0, e, 0)), 1, 0);
    R gz = F(sp_sample)(&c, F(add)(hp, F(mk)(0, 0, e)), 1, 0) - F(sp_sample)(&c, F(sub)(hp, F(mk)(0, 0, e)), 1, 0);
    V3 g = F(mk)(gx, gy, gz);
    R len = F(length)(g);
    V3 n = len < (R)1e-10f ? F(mk)(0, 1, 0) : F(mk)(g.x / len, g.y / len, g.z / len);
    uint32_t sl, vx, cell;
    int64_t vi = F(sp_nearest)(&c, hp, &sl, &vx, &cell);
    /* secondary material and blend weight of the nearest voxel (:570-596),
       from the aux plane (secondary | blend << 8) when the scene has one */
    uint32_t sec = 0, blend = 0;
    if (s->aux && vi >= 0) {
        sec = s->aux[vi] & 0x3Fu;
        blend = (uint32_t)(s->aux[vi] >> 8) & 0xFFu;
    }
    out->brick_slot = sl;
    out->voxel_index = vx;
    out->cell_index = cell;
    R sdf_at = F(sp_sample)(&c, hp, 0, 0);
    out->flags = c.flags;
    out->hit = 1;
    out->t = (double)r.t;
    out->pos[0] = (double)hp.x;
    out->pos[1] = (double)hp.y;
    out->pos[2] = (double)hp.z;
    out->normal[0] = (double)n.x;
    out->normal[1] = (double)n.y;
    out->normal[2] = (double)n.z;
    out->material_id = mat;
    out->packed_material = (mat & 0x3Fu) | (sec << 6) | (blend << 12);
    R mx, my;
    F(motion)(cam, w, h, px, py, hp, &mx, &my);
    out->motion[0] = (double)mx;
    out->motion[1] = (double)my;
    out->sdf_at_hit = (double)sdf_at;
    if (ext->shadow_enable) {
        /* The shading pass reads the normal back from the rgba16f target. */
        V3 nq = F(mk)((R)orc_f16_to_f32(orc_f32_to_f16((float)n.x)), (R)orc_f16_to_f32(orc_f32_to_f16((float)n.y)),
                      (R)orc_f16_to_f32(orc_f32_to_f16((float)n.z)));
        V3 L = F(mk)((R)ext->light_dir[0], (R)ext->light_dir[1], (R)ext->light_dir[2]);
        R bias = (R)0.08f;
        V3 so = F(add)(F(add)(hp, F(muls)(nq, bias)), F(muls)(L, bias * (R)0.5f));
        R maxd = ext->shadow_max_dist > 0.0f ? (R)ext->shadow_max_dist : (R)25.0f;
        R k = ext->shadow_k > 0.0f ? (R)ext->shadow_k : (R)16.0f;
        F(sctx) cs;
        F(sp_init)(&cs, s);
        out->shadow = (double)F(sp_shadow)(&cs, so, L, maxd, k, &out->shadow_steps);
        out->shadow_evals = cs.evals;
    }
}

/* Sphere casts (rkfm_sparse_sphere_cast's definition, include/rkf_march.h):
 * L(p) from the free-space structures, D = L when L >= 16 vs, else
 * max(stepping sample, L); hit at D - r < 1e-3. */
static R F(sp_free_bound)(const F(sctx) *c, V3 p) {
    const orc_sparse_scene *s = c->s;
    R vs = (R)s->voxel_size, be = vs * (R)8.0f;
    V3 gs = F(mk)((R)c->nb[0] * be, (R)c->nb[1] * be, (R)c->nb[2] * be);
    V3 gp = F(sub)(p, F(mk)((R)s->origin[0], (R)s->origin[1], (R)s->origin[2]));
    if (gp.x < (R)0 || gp.y < (R)0 || gp.z < (R)0 || gp.x >= gs.x || gp.y >= gs.y || gp.z >= gs.z) {
        R dx = FMAX(FMAX(-gp.x, gp.x - gs.x), (R)0), dy = FMAX(FMAX(-gp.y, gp.y - gs.y), (R)0),
          dz = FMAX(FMAX(-gp.z, gp.z - gs.z), (R)0);
        return SQRT(dx * dx + dy * dy + dz * dz);
    }
    R ivs = (R)1.0f / vs;
    R x[3] = {gp.x * ivs, gp.y * ivs, gp.z * ivs};
    uint32_t b[3];
    for (int a = 0; a < 3; a++) b[a] = (uint32_t)F(clampi)((int32_t)FLOOR(x[a]), 0, c->tot[a] - 1) >> 3;
    uint64_t cell = (uint64_t)(b[0] >> 3) + (uint64_t)(b[1] >> 3) * s->cells[0] +
                    (uint64_t)(b[2] >> 3) * s->cells[0] * s->cells[1];
    uint32_t e = 15u;
    for (uint32_t o = 0; o < 8u; o++) {
        uint32_t v = (s->oct[cell] >> (4u * o)) & 15u;
        e = v < e ? v : e;
    }
    R unit = (R)64.0f;
    uint32_t base[3] = {b[0] >> 3, b[1] >> 3, b[2] >> 3};
    if (e == 0u) {
        if (s->free_l2[cell]) return 0;
        uint32_t lb = (b[0] & 7u) + (b[1] & 7u) * 8u + (b[2] & 7u) * 64u;
        uint32_t w = s->bext[cell * 512u + lb];
        e = 15u;
        for (uint32_t o = 0; o < 8u; o++) {
            uint32_t v = (w >> (4u * o)) & 15u;
            e = v < e ? v : e;
        }
        if (e == 0u) return 0;
        unit = (R)8.0f;
        for (int a = 0; a < 3; a++) base[a] = b[a];
    }
    R m = K_MAX_FLOAT;
    for (int a = 0; a < 3; a++) {
        R lo = ((R)base[a] - (R)(e - 1u)) * unit, hi = ((R)base[a] + (R)e) * unit;
        m = FMIN(m, FMIN(x[a] - lo, hi - x[a]));
    }
    return FMAX(m, (R)0) * vs;
}

static void F(sp_cast)(const orc_sparse_scene *s, const rkfm_sphere_cast *q, rkfm_cast_result *out) {
    F(sctx) c;
    F(sp_init)(&c, s);
    V3 o = F(mk)((R)q->origin[0], (R)q->origin[1], (R)q->origin[2]);
    V3 dir = F(mk)((R)q->dir[0], (R)q->dir[1], (R)q->dir[2]);
    R r = (R)q->radius, vs = (R)s->voxel_size, t = 0;
    memset(out, 0, sizeof(*out));
    out->t = q->max_t;
    for (uint32_t i = 0; i < K_HARD_MAX; i++) {
        if (t > (R)q->max_t) break;
        out->steps++;
        V3 p = F(add)(o, F(muls)(dir, t));
        R L = F(sp_free_bound)(&c, p);
        R D = L >= vs * (R)16.0f ? L : FMAX(F(sp_sample)(&c, p, 0, 1), L);
        out->distance = (float)D;
        if (D - r < (R)1e-3f) {
            R e = vs * (R)2.0f;
            R gx = F(sp_sample)(&c, F(add)(p, F(mk)(e, 0, 0)), 1, 0) - F(sp_sample)(&c, F(sub)(p, F(mk)(e, 0, 0)), 1, 0);
            R gy = F(sp_sample)(&c, F(add)(p, F(mk)(0, e, 0)), 1, 0) - F(sp_sample)(&c, F(sub)(p, F(mk)(0, e, 0)), 1, 0);
            R gz = F(sp_sample)(&c, F(add)(p, F(mk)(0, 0, e)), 1, 0) - F(sp_sample)(&c, F(sub)(p, F(mk)(0, 0, e)), 1, 0);
            R len = F(length)(F(mk)(gx, gy, gz));
            V3 n = len < (R)1e-10f ? F(mk)(0, 1, 0) : F(mk)(gx / len, gy / len, gz / len);
            out->t = (float)t;
            out->hit = 1;
            out->normal[0] = (float)n.x;
            out->normal[1] = (float)n.y;
            out->normal[2] = (float)n.z;
            return;
        }
        t += FMAX(D - r, K_MIN_STEP);
    }
}

/* Point query: SdfQueryable::evaluate + gradient_central_diff
 * (sdf_collision.rs:38-56; eps 0.01, normalize_or_zero = v * (1/len) if that
 * reciprocal is finite and > 0). */
static void F(sp_query)(const orc_sparse_scene *s, const float *xyz, rkfm_point_result *out) {
    F(sctx) c;
    F(sp_init)(&c, s);
    V3 p = F(mk)((R)xyz[0], (R)xyz[1], (R)xyz[2]);
    R d = F(sp_sample)(&c, p, 0, 0);
    R e = (R)0.01f;
    R dx = F(sp_sample)(&c, F(add)(p, F(mk)(e, 0, 0)), 0, 0) - F(sp_sample)(&c, F(sub)(p, F(mk)(e, 0, 0)), 0, 0);
    R dy = F(sp_sample)(&c, F(add)(p, F(mk)(0, e, 0)), 0, 0) - F(sp_sample)(&c, F(sub)(p, F(mk)(0, e, 0)), 0, 0);
    R dz = F(sp_sample)(&c, F(add)(p, F(mk)(0, 0, e)), 0, 0) - F(sp_sample)(&c, F(sub)(p, F(mk)(0, 0, e)), 0, 0);
    R len = SQRT(dx * dx + dy * dy + dz * dz);
    R rcp = (R)1 / len;
    out->distance = (float)d;
    out->material = F(sp_material)(&c, p);
    out->inside = d < (R)0 ? 1u : 0u;
    if (ISFINITE(rcp) && rcp > (R)0) {
        out->gradient[0] = (float)(dx * rcp);
        out->gradient[1] = (float)(dy * rcp);
        out->gradient[2] = (float)(dz * rcp);
    } else {
        out->gradient[0] = out->gradient[1] = out->gradient[2] = 0.0f;
    }
}

#undef V3

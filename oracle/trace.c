/*
 * trace.c — instantiates trace_impl.h as the FP32 twin and the FP64 oracle,
 * and restates the march's producers (coarse field, tile cull). Test
 * infrastructure only (see oracle.h). Build with -ffp-contract=off.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"

#define R float
#define SFX _f32
#define SQRT sqrtf
#define FLOOR floorf
#define FABS fabsf
#define FMIN fminf
#define FMAX fmaxf
#define ISFINITE isfinite
#include "trace_impl.h"
#undef R
#undef SFX
#undef SQRT
#undef FLOOR
#undef FABS
#undef FMIN
#undef FMAX

#define R double
#define SFX _f64
#define SQRT sqrt
#define FLOOR floor
#define FABS fabs
#define FMIN fmin
#define FMAX fmax
#include "trace_impl.h"
#undef R
#undef SFX

void orc_ref_trace(const orc_ref_scene *s, int prec, uint64_t p0, uint64_t p1, orc_ray_result *out) {
    for (uint64_t p = p0; p < p1; p++) {
        if (prec == 64) ref_pixel_f64(s, p, out + (p - p0));
        else ref_pixel_f32(s, p, out + (p - p0));
    }
}

void orc_ref_eval_object(const orc_ref_scene *s, int prec, uint32_t oi, const double p[3], double *dist,
                         uint32_t *mat) {
    if (prec == 64) {
        rctx_f64 c = {s, 0, 0, 0};
        *dist = eval_object_f64(&c, mk_f64(p[0], p[1], p[2]), oi, mat);
    } else {
        rctx_f32 c = {s, 0, 0, 0};
        *dist = (double)eval_object_f32(&c, mk_f32((float)p[0], (float)p[1], (float)p[2]), oi, mat);
    }
}

/* ---- reference scene -> sparse scene (rkfm_sparse_from_objects) ---- */
static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : x > y;
}

/* min over the cell's candidate objects (ascending index) of evaluate_object;
 * *win = first object reaching it (UINT32_MAX if none), *mat its material. */
static float objects_field_f32(const orc_ref_scene *s, const uint32_t *objs, uint32_t n, v3_f32 p, uint32_t *win,
                               uint32_t *mat) {
    rctx_f32 c = {s, 0, 0, 0};
    float best = INFINITY;
    *win = UINT32_MAX;
    *mat = 0;
    for (uint32_t k = 0; k < n; k++) {
        uint32_t m;
        float d = eval_object_f32(&c, p, objs[k], &m);
        if (d < best) {
            best = d;
            *win = objs[k];
            *mat = m;
        }
    }
    return best;
}

int orc_sparse_from_objects(const orc_ref_scene *s, float vs, const float bmin[3], const float bmax[3],
                            uint32_t cells[3], uint32_t *n_blocks, uint64_t *n_bricks, uint32_t *l2, uint32_t *l1,
                            uint16_t *dist, uint8_t *mat, uint16_t *aux) {
    float bws = vs * 8.0f, cws = bws * 8.0f;
    for (int a = 0; a < 3; a++) {
        double cc = ceil((double)(bmax[a] - bmin[a]) / (double)cws);
        cells[a] = cc < 1.0 ? 1u : (uint32_t)cc;
    }
    uint64_t ncell = (uint64_t)cells[0] * cells[1] * cells[2];
    uint32_t nobj = s->su.num_objects;
    double M = 3.0 * (double)bws;
    /* (cell << 32 | object) pairs, sorted: the candidate lists */
    uint64_t cap = 1024, np = 0;
    uint64_t *pairs = (uint64_t *)malloc(cap * 8);
    for (uint32_t j = 0; j < nobj; j++) {
        const rkf_gpu_object *o = &s->objects[j];
        if (o->sdf_type == RKF_SDF_TYPE_NONE) continue;
        int64_t lo[3], hi[3];
        for (int a = 0; a < 3; a++) {
            lo[a] = (int64_t)floor(((double)o->aabb_min[a] - M - (double)bmin[a]) / (double)cws);
            hi[a] = (int64_t)floor(((double)o->aabb_max[a] + M - (double)bmin[a]) / (double)cws);
            if (lo[a] < 0) lo[a] = 0;
            if (hi[a] > (int64_t)cells[a] - 1) hi[a] = (int64_t)cells[a] - 1;
        }
        for (int64_t z = lo[2]; z <= hi[2]; z++)
            for (int64_t y = lo[1]; y <= hi[1]; y++)
                for (int64_t x = lo[0]; x <= hi[0]; x++) {
                    if (np == cap) { cap *= 2; pairs = (uint64_t *)realloc(pairs, cap * 8); }
                    pairs[np++] = (((uint64_t)x + (uint64_t)y * cells[0] + (uint64_t)z * cells[0] * cells[1]) << 32) | j;
                }
    }
    qsort(pairs, np, 8, cmp_u64);
    uint32_t *objs = (uint32_t *)malloc((np + 1) * 4);
    uint32_t blocks = 0;
    uint64_t bricks = 0;
    if (l2) for (uint64_t i = 0; i < ncell; i++) l2[i] = RKF_EMPTY_SLOT;
    float hb = bws * 0.5f, half = vs * 0.5f;
    for (uint64_t i = 0; i < np;) {
        uint32_t cell = (uint32_t)(pairs[i] >> 32), n = 0;
        while (i < np && (uint32_t)(pairs[i] >> 32) == cell) objs[n++] = (uint32_t)pairs[i++];
        uint32_t qx = cell % cells[0], qy = (cell / cells[0]) % cells[1], qz = cell / (cells[0] * cells[1]);
        uint32_t stored = 0;
        for (uint32_t lb = 0; lb < 512u; lb++) {
            uint32_t gbx = qx * 8u + (lb & 7u), gby = qy * 8u + ((lb >> 3) & 7u), gbz = qz * 8u + (lb >> 6);
            v3_f32 c = mk_f32((bmin[0] + (float)gbx * bws) + hb, (bmin[1] + (float)gby * bws) + hb,
                              (bmin[2] + (float)gbz * bws) + hb);
            uint32_t win, m;
            float d = objects_field_f32(s, objs, n, c, &win, &m);
            int need = fabsf(d) < bws * 1.8f;
            if (l2 && need) {
                if (!stored) {
                    l2[cell] = blocks;
                    for (uint32_t k = 0; k < 512u; k++) l1[(uint64_t)blocks * 512u + k] = RKF_EMPTY_SLOT;
                }
                uint64_t slot = bricks + stored;
                l1[(uint64_t)blocks * 512u + lb] = (uint32_t)slot;
                for (uint32_t v = 0; v < 512u; v++) {
                    uint32_t vx = v & 7u, vy = (v >> 3) & 7u, vz = v >> 6;
                    v3_f32 p = mk_f32((bmin[0] + (float)gbx * bws) + ((float)vx * vs + half),
                                      (bmin[1] + (float)gby * bws) + ((float)vy * vs + half),
                                      (bmin[2] + (float)gbz * bws) + ((float)vz * vs + half));
                    uint32_t w, mm;
                    float dv = objects_field_f32(s, objs, n, p, &w, &mm);
                    uint64_t idx = slot * 512u + v;
                    dist[idx] = orc_f32_to_f16(dv);
                    mat[idx] = (uint8_t)(mm & 0x3Fu);
                    uint32_t a = 0;
                    if (w != UINT32_MAX && s->objects[w].sdf_type == RKF_SDF_TYPE_VOXELIZED) {
                        const rkf_gpu_object *o = &s->objects[w];
                        uint32_t sl, vo;
                        uint64_t vi = nearest_voxel_f32(s, xform_point_f32(o->inverse_world, p), o, &sl, &vo);
                        if (vi != UINT64_MAX)
                            a = ((pool_word0_f32(s, vi) >> 22) & 0x3Fu) | (((pool_word1_f32(s, vi) >> 24) & 0xFFu) << 8);
                    }
                    if (aux) aux[idx] = (uint16_t)a;
                }
            }
            stored += need;
        }
        if (stored) {
            blocks++;
            bricks += stored;
        }
    }
    free(pairs);
    free(objs);
    *n_blocks = blocks;
    *n_bricks = bricks;
    return 0;
}

void orc_sparse_cast(const orc_sparse_scene *s, const rkfm_sphere_cast *casts, uint64_t n, int prec,
                     rkfm_cast_result *out) {
    for (uint64_t i = 0; i < n; i++) {
        if (prec == 64) sp_cast_f64(s, &casts[i], &out[i]);
        else sp_cast_f32(s, &casts[i], &out[i]);
    }
}

int orc_ref_trace_parallel(const orc_ref_scene *s, orc_ray_result *out, int threads) {
    uint64_t n = (uint64_t)s->width * s->height;
    int used = 1;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel
    {
#pragma omp single
        used = omp_get_num_threads();
#pragma omp for schedule(dynamic, 64)
        for (int64_t p = 0; p < (int64_t)n; p++) ref_pixel_f32(s, (uint64_t)p, out + p);
    }
#else
    (void)threads;
    for (uint64_t p = 0; p < n; p++) ref_pixel_f32(s, p, out + p);
#endif
    return used;
}

uint64_t orc_sparse_n_bricks_global(const orc_sparse_scene *s) {
    return (uint64_t)s->cells[0] * 8u * s->cells[1] * 8u * s->cells[2] * 8u;
}

static int brick_empty(const orc_sparse_scene *s, uint32_t x, uint32_t y, uint32_t z) {
    uint64_t cell = (uint64_t)(x >> 3) + (uint64_t)(y >> 3) * s->cells[0] +
                    (uint64_t)(z >> 3) * s->cells[0] * s->cells[1];
    uint32_t blk = s->l2[cell];
    if (blk == RKF_EMPTY_SLOT) return 1;
    return s->l1[(uint64_t)blk * 512u + (x & 7u) + (y & 7u) * 8u + (z & 7u) * 64u] == RKF_EMPTY_SLOT;
}

/* Brick b is free when every brick of {b..b+1}^3 (clamped to the grid) is EMPTY. */
int orc_sparse_brick_free(const orc_sparse_scene *s, uint32_t bx, uint32_t by, uint32_t bz) {
    uint32_t nb[3] = {s->cells[0] * 8u, s->cells[1] * 8u, s->cells[2] * 8u};
    for (uint32_t dz = 0; dz < 2; dz++)
        for (uint32_t dy = 0; dy < 2; dy++)
            for (uint32_t dx = 0; dx < 2; dx++) {
                uint32_t x = bx + dx < nb[0] ? bx + dx : nb[0] - 1;
                uint32_t y = by + dy < nb[1] ? by + dy : nb[1] - 1;
                uint32_t z = bz + dz < nb[2] ? bz + dz : nb[2] - 1;
                if (!brick_empty(s, x, y, z)) return 0;
            }
    return 1;
}

/* free_l2[cell] = every brick b of the cell is free, i.e. no allocated brick
 * b' lies in [8C, 8C+8]^3 (brick coordinates, clamped): each allocated brick
 * b' clears the cells C with 8C <= b' <= 8C+8 per axis. */
void orc_sparse_free_l2(const orc_sparse_scene *s, uint8_t *free_l2) {
    uint64_t nc = (uint64_t)s->cells[0] * s->cells[1] * s->cells[2];
    for (uint64_t c = 0; c < nc; c++) free_l2[c] = 1;
    for (uint32_t qz = 0; qz < s->cells[2]; qz++)
        for (uint32_t qy = 0; qy < s->cells[1]; qy++)
            for (uint32_t qx = 0; qx < s->cells[0]; qx++) {
                uint64_t cell = (uint64_t)qx + (uint64_t)qy * s->cells[0] + (uint64_t)qz * s->cells[0] * s->cells[1];
                uint32_t blk = s->l2[cell];
                if (blk == RKF_EMPTY_SLOT) continue;
                for (uint32_t lb = 0; lb < 512; lb++) {
                    if (s->l1[(uint64_t)blk * 512u + lb] == RKF_EMPTY_SLOT) continue;
                    uint32_t b[3] = {qx * 8u + (lb & 7u), qy * 8u + ((lb >> 3) & 7u), qz * 8u + (lb >> 6)};
                    uint32_t lo[3], hi[3];
                    for (int a = 0; a < 3; a++) {
                        /* cells C with 8C <= b <= 8C + 8 */
                        hi[a] = b[a] >> 3;
                        lo[a] = (b[a] >= 8u && (b[a] & 7u) == 0) ? hi[a] - 1 : hi[a];
                    }
                    for (uint32_t z = lo[2]; z <= hi[2]; z++)
                        for (uint32_t y = lo[1]; y <= hi[1]; y++)
                            for (uint32_t x = lo[0]; x <= hi[0]; x++)
                                free_l2[(uint64_t)x + (uint64_t)y * s->cells[0] +
                                        (uint64_t)z * s->cells[0] * s->cells[1]] = 0;
                }
            }
}

void orc_sparse_oct(const orc_sparse_scene *s, const uint8_t *free_l2, uint32_t *oct) {
    const int64_t cx = s->cells[0], cy = s->cells[1], cz = s->cells[2];
    const uint32_t cap = 15;
    for (int64_t i = 0; i < cx * cy * cz; i++) oct[i] = 0;
    for (int o = 0; o < 8; o++) {
        const int sx = (o & 1) ? -1 : 1, sy = (o & 2) ? -1 : 1, sz = (o & 4) ? -1 : 1;
        /* visit cells so that the neighbours further along o come first */
        for (int64_t kz = 0; kz < cz; kz++)
            for (int64_t ky = 0; ky < cy; ky++)
                for (int64_t kx = 0; kx < cx; kx++) {
                    const int64_t x = sx > 0 ? cx - 1 - kx : kx;
                    const int64_t y = sy > 0 ? cy - 1 - ky : ky;
                    const int64_t z = sz > 0 ? cz - 1 - kz : kz;
                    const int64_t i = x + y * cx + z * cx * cy;
                    if (!free_l2[i]) continue; /* E = 0 */
                    uint32_t m = cap;
                    for (int k = 1; k < 8; k++) {
                        const int64_t xx = x + (k & 1) * sx, yy = y + ((k >> 1) & 1) * sy, zz = z + (k >> 2) * sz;
                        uint32_t v = cap; /* out of the grid: free */
                        if (xx >= 0 && yy >= 0 && zz >= 0 && xx < cx && yy < cy && zz < cz)
                            v = (oct[xx + yy * cx + zz * cx * cy] >> (4 * o)) & 15u;
                        if (v < m) m = v;
                    }
                    const uint32_t e = m + 1 < cap ? m + 1 : cap;
                    oct[i] |= e << (4 * o);
                }
    }
}

void orc_sparse_bext(const orc_sparse_scene *s, const uint8_t *free_l2, const uint32_t *oct, uint32_t *out) {
    const int64_t cx = s->cells[0], cy = s->cells[1], cz = s->cells[2];
    const int64_t nx = cx * 8, ny = cy * 8, nz = cz * 8;
    const uint32_t cap = 15;
    for (int64_t i = 0; i < cx * cy * cz * 512; i++) out[i] = 0;
    for (int o = 0; o < 8; o++) {
        const int sg[3] = {(o & 1) ? -1 : 1, (o & 2) ? -1 : 1, (o & 4) ? -1 : 1};
        for (int64_t kz = 0; kz < nz; kz++)
            for (int64_t ky = 0; ky < ny; ky++)
                for (int64_t kx = 0; kx < nx; kx++) {
                    const int64_t b[3] = {sg[0] > 0 ? nx - 1 - kx : kx, sg[1] > 0 ? ny - 1 - ky : ky,
                                          sg[2] > 0 ? nz - 1 - kz : kz};
                    const int64_t cell = (b[0] >> 3) + (b[1] >> 3) * cx + (b[2] >> 3) * cx * cy;
                    if (free_l2[cell]) continue;
                    if (!orc_sparse_brick_free(s, (uint32_t)b[0], (uint32_t)b[1], (uint32_t)b[2])) continue; /* 0 */
                    uint32_t m = cap;
                    for (int k = 1; k < 8; k++) {
                        const int64_t n[3] = {b[0] + (k & 1) * sg[0], b[1] + ((k >> 1) & 1) * sg[1],
                                              b[2] + (k >> 2) * sg[2]};
                        uint32_t v = cap;
                        if (n[0] >= 0 && n[1] >= 0 && n[2] >= 0 && n[0] < nx && n[1] < ny && n[2] < nz) {
                            const int64_t c2 = (n[0] >> 3) + (n[1] >> 3) * cx + (n[2] >> 3) * cx * cy;
                            if (!free_l2[c2]) {
                                const int64_t lb = (n[0] & 7) + (n[1] & 7) * 8 + (n[2] & 7) * 64;
                                v = (out[c2 * 512 + lb] >> (4 * o)) & 15u;
                            } else {
                                const int64_t e = (oct[c2] >> (4 * o)) & 15u;
                                const int64_t cc[3] = {n[0] >> 3, n[1] >> 3, n[2] >> 3};
                                int64_t mm = cap;
                                for (int a = 0; a < 3; a++) {
                                    int64_t d = sg[a] < 0 ? n[a] - (cc[a] - e + 1) * 8 + 1 : (cc[a] + e) * 8 - n[a];
                                    if (d < mm) mm = d;
                                }
                                v = (uint32_t)mm;
                            }
                        }
                        if (v < m) m = v;
                    }
                    const uint32_t e = m + 1 < cap ? m + 1 : cap;
                    out[cell * 512 + (b[0] & 7) + (b[1] & 7) * 8 + (b[2] & 7) * 64] |= e << (4 * o);
                }
    }
}

/* ---- contacts: sdf_collision.rs:79-148 (sample points), :188-225 ---- */
static int body_points(const rkfm_body *b, float pts[16][3]) {
    int n = 0;
#define PUSH(X, Y, Z) do { pts[n][0] = (X); pts[n][1] = (Y); pts[n][2] = (Z); n++; } while (0)
    if (b->shape == RKFM_BODY_SPHERE) {
        float r = b->dims[0], d3 = r / sqrtf(3.0f);
        PUSH(r, 0.0f, 0.0f); PUSH(-r, 0.0f, 0.0f); PUSH(0.0f, r, 0.0f);
        PUSH(0.0f, -r, 0.0f); PUSH(0.0f, 0.0f, r); PUSH(0.0f, 0.0f, -r);
        PUSH(d3, d3, d3); PUSH(d3, d3, -d3); PUSH(d3, -d3, d3); PUSH(d3, -d3, -d3);
        PUSH(-d3, d3, d3); PUSH(-d3, d3, -d3); PUSH(-d3, -d3, d3); PUSH(-d3, -d3, -d3);
    } else if (b->shape == RKFM_BODY_BOX) {
        const float sgn[2] = {-1.0f, 1.0f};
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++)
                for (int k = 0; k < 2; k++) PUSH(b->dims[0] * sgn[i], b->dims[1] * sgn[j], b->dims[2] * sgn[k]);
        PUSH(b->dims[0], 0.0f, 0.0f); PUSH(-b->dims[0], 0.0f, 0.0f); PUSH(0.0f, b->dims[1], 0.0f);
        PUSH(0.0f, -b->dims[1], 0.0f); PUSH(0.0f, 0.0f, b->dims[2]); PUSH(0.0f, 0.0f, -b->dims[2]);
    } else if (b->shape == RKFM_BODY_CAPSULE) {
        float hh = b->dims[0], r = b->dims[1];
        PUSH(0.0f, hh + r, 0.0f); PUSH(r, hh, 0.0f); PUSH(-r, hh, 0.0f); PUSH(0.0f, hh, r); PUSH(0.0f, hh, -r);
        PUSH(0.0f, -(hh + r), 0.0f); PUSH(r, -hh, 0.0f); PUSH(-r, -hh, 0.0f); PUSH(0.0f, -hh, r); PUSH(0.0f, -hh, -r);
        PUSH(r, 0.0f, 0.0f); PUSH(-r, 0.0f, 0.0f);
    }
#undef PUSH
    return n;
}

void orc_sparse_contacts(const orc_sparse_scene *s, const rkfm_body *bodies, uint32_t n, rkfm_contact *out,
                         uint32_t *counts) {
    for (uint32_t i = 0; i < n; i++) {
        const rkfm_body *b = &bodies[i];
        float pts[16][3];
        int np = body_points(b, pts);
        sctx_f32 c;
        sp_init_f32(&c, s);
        rkfm_contact list[16];
        uint32_t m = 0;
        for (int k = 0; k < np; k++) {
            /* glam Quat * Vec3 in scalar form */
            const float *q = b->rotation;
            float bx = q[0], by = q[1], bz = q[2], w = q[3];
            float b2 = bx * bx + by * by + bz * bz, sc = w * w - b2;
            float *v = pts[k];
            float d2 = (v[0] * bx + v[1] * by + v[2] * bz) * 2.0f, w2 = w * 2.0f;
            float cr[3] = {by * v[2] - bz * v[1], bz * v[0] - bx * v[2], bx * v[1] - by * v[0]};
            v3_f32 wp = mk_f32(b->position[0] + ((v[0] * sc + bx * d2) + cr[0] * w2),
                               b->position[1] + ((v[1] * sc + by * d2) + cr[1] * w2),
                               b->position[2] + ((v[2] * sc + bz * d2) + cr[2] * w2));
            float d = sp_sample_f32(&c, wp, 0, 0);
            if (!(d < b->threshold)) continue;
            const float e = 0.01f;
            float dx = sp_sample_f32(&c, add_f32(wp, mk_f32(e, 0, 0)), 0, 0) - sp_sample_f32(&c, sub_f32(wp, mk_f32(e, 0, 0)), 0, 0);
            float dy = sp_sample_f32(&c, add_f32(wp, mk_f32(0, e, 0)), 0, 0) - sp_sample_f32(&c, sub_f32(wp, mk_f32(0, e, 0)), 0, 0);
            float dz = sp_sample_f32(&c, add_f32(wp, mk_f32(0, 0, e)), 0, 0) - sp_sample_f32(&c, sub_f32(wp, mk_f32(0, 0, e)), 0, 0);
            float rcp = 1.0f / sqrtf(dx * dx + dy * dy + dz * dz);
            rkfm_contact ct;
            ct.position[0] = wp.x; ct.position[1] = wp.y; ct.position[2] = wp.z;
            if (isfinite(rcp) && rcp > 0.0f) {
                ct.normal[0] = dx * rcp; ct.normal[1] = dy * rcp; ct.normal[2] = dz * rcp;
            } else {
                ct.normal[0] = ct.normal[1] = ct.normal[2] = 0.0f;
            }
            ct.penetration = b->threshold - d;
            ct.sample = (uint32_t)k;
            /* stable insertion: deepest first */
            uint32_t at = m;
            while (at > 0 && list[at - 1].penetration < ct.penetration) {
                list[at] = list[at - 1];
                at--;
            }
            list[at] = ct;
            m++;
        }
        memcpy(out + (uint64_t)i * RKFM_MAX_CONTACTS, list, m * sizeof(rkfm_contact));
        counts[i] = m;
    }
}

/* ---- edits: cpu_apply.rs:28-281, edit_op.rs:86-140 ---- */
static float e_len3(float x, float y, float z) { return sqrtf(x * x + y * y + z * z); }

/* glam's Quat * Vec3 in scalar form: v (w^2 - b.b) + b (2 v.b) + (b x v)(2 w). */
static void e_rotate(const float q[4], const float v[3], float o[3]) {
    float bx = q[0], by = q[1], bz = q[2], w = q[3];
    float b2 = bx * bx + by * by + bz * bz;
    float sc = w * w - b2;
    float d2 = (v[0] * bx + v[1] * by + v[2] * bz) * 2.0f;
    float c[3] = {by * v[2] - bz * v[1], bz * v[0] - bx * v[2], bx * v[1] - by * v[0]};
    float w2 = w * 2.0f;
    o[0] = (v[0] * sc + bx * d2) + c[0] * w2;
    o[1] = (v[1] * sc + by * d2) + c[1] * w2;
    o[2] = (v[2] * sc + bz * d2) + c[2] * w2;
}

float orc_edit_shape(uint32_t shape, const float dims[3], const float p[3]) {
    switch (shape) {
    case RKFM_SHAPE_SPHERE: return e_len3(p[0], p[1], p[2]) - dims[0];
    case RKFM_SHAPE_BOX: {
        float q[3] = {fabsf(p[0]) - dims[0], fabsf(p[1]) - dims[1], fabsf(p[2]) - dims[2]};
        float outside = e_len3(fmaxf(q[0], 0.0f), fmaxf(q[1], 0.0f), fmaxf(q[2], 0.0f));
        float inside = fminf(fmaxf(q[0], fmaxf(q[1], q[2])), 0.0f);
        return outside + inside;
    }
    case RKFM_SHAPE_CAPSULE: {
        float a[3] = {0.0f, -dims[1], 0.0f}, b[3] = {0.0f, dims[1], 0.0f};
        float pa[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
        float ba[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        float h = (pa[0] * ba[0] + pa[1] * ba[1] + pa[2] * ba[2]) / (ba[0] * ba[0] + ba[1] * ba[1] + ba[2] * ba[2]);
        h = h < 0.0f ? 0.0f : (h > 1.0f ? 1.0f : h);
        return e_len3(pa[0] - ba[0] * h, pa[1] - ba[1] * h, pa[2] - ba[2] * h) - dims[0];
    }
    case RKFM_SHAPE_CYLINDER: {
        float d_radial = e_len3(p[0], 0.0f, p[2]) - dims[0];
        float d_height = fabsf(p[1]) - dims[1];
        float ox = fmaxf(d_radial, 0.0f), oy = fmaxf(d_height, 0.0f);
        return sqrtf(ox * ox + oy * oy) + fminf(fmaxf(d_radial, d_height), 0.0f);
    }
    case RKFM_SHAPE_TORUS: {
        float qx = e_len3(p[0], 0.0f, p[2]) - dims[0];
        return sqrtf(qx * qx + p[1] * p[1]) - dims[1];
    }
    default: return p[1] - dims[0];
    }
}

float orc_edit_falloff(uint32_t curve, float distance, float radius) {
    if (radius <= 0.0f || distance >= radius) return 0.0f;
    float t = distance / radius;
    switch (curve) {
    case RKFM_FALLOFF_LINEAR: return 1.0f - t;
    case RKFM_FALLOFF_SHARP: { float f = 1.0f - t; return f * f * f; }
    default: return 1.0f - t * t * (3.0f - 2.0f * t);
    }
}

static float e_smin(float a, float b, float k) {
    float h = fmaxf(k - fabsf(a - b), 0.0f) / k;
    return fminf(a, b) - h * h * k * 0.25f;
}

void orc_sparse_edit(const orc_sparse_scene *s, uint16_t *dist, uint8_t *mat, const rkfm_edit *ops,
                     uint32_t n_ops, uint8_t *modified) {
    const float vs = s->voxel_size, bs = vs * 8.0f;
    const int32_t nb[3] = {(int32_t)s->cells[0] * 8, (int32_t)s->cells[1] * 8, (int32_t)s->cells[2] * 8};
    for (uint32_t k = 0; k < n_ops; k++) {
        const rkfm_edit *op = &ops[k];
        const float max_dim = fmaxf(fmaxf(op->dimensions[0], op->dimensions[1]), op->dimensions[2]);
        const float extent = max_dim + op->blend_k;
        int32_t bmin[3], bmax[3];
        for (int a = 0; a < 3; a++) {
            float lo = floorf(((op->position[a] - extent) - s->origin[a]) / bs);
            float hi = ceilf(((op->position[a] + extent) - s->origin[a]) / bs - 0.001f);
            int32_t l = (int32_t)lo, h = (int32_t)hi;
            bmin[a] = l < 0 ? 0 : l;
            h = h < 0 ? 0 : h;
            bmax[a] = h < nb[a] - 1 ? h : nb[a] - 1;
        }
        const float iq[4] = {-op->rotation[0], -op->rotation[1], -op->rotation[2], op->rotation[3]};
        for (int32_t bz = bmin[2]; bz <= bmax[2]; bz++)
            for (int32_t by = bmin[1]; by <= bmax[1]; by++)
                for (int32_t bx = bmin[0]; bx <= bmax[0]; bx++) {
                    uint64_t cell = (uint64_t)(bx >> 3) + (uint64_t)(by >> 3) * s->cells[0] +
                                    (uint64_t)(bz >> 3) * s->cells[0] * s->cells[1];
                    uint32_t blk = s->l2[cell];
                    if (blk == RKF_EMPTY_SLOT) continue;
                    uint32_t slot = s->l1[(uint64_t)blk * 512u + (bx & 7) + (by & 7) * 8 + (bz & 7) * 64];
                    if (slot == RKF_EMPTY_SLOT || slot == RKF_INTERIOR_SLOT) continue;
                    const float bmn[3] = {s->origin[0] + (float)bx * bs, s->origin[1] + (float)by * bs,
                                          s->origin[2] + (float)bz * bs};
                    int any = 0;
                    for (uint32_t vz = 0; vz < 8; vz++)
                        for (uint32_t vy = 0; vy < 8; vy++)
                            for (uint32_t vx = 0; vx < 8; vx++) {
                                const float vp[3] = {bmn[0] + (float)vx * vs + vs * 0.5f,
                                                     bmn[1] + (float)vy * vs + vs * 0.5f,
                                                     bmn[2] + (float)vz * vs + vs * 0.5f};
                                const float rel[3] = {vp[0] - op->position[0], vp[1] - op->position[1],
                                                      vp[2] - op->position[2]};
                                float loc[3];
                                e_rotate(iq, rel, loc);
                                float shape_d = orc_edit_shape(op->shape_type, op->dimensions, loc);
                                float w = orc_edit_falloff(op->falloff, e_len3(rel[0], rel[1], rel[2]),
                                                           max_dim + op->blend_k);
                                if (w <= 0.0f) continue;
                                uint64_t vi = (uint64_t)slot * 512u + vx + vy * 8u + vz * 64u;
                                float existing = orc_f16_to_f32(dist[vi]);
                                uint32_t m = mat[vi];
                                float strength = op->strength * w;
                                float thr = vs * 0.1f;
                                float nd = existing;
                                uint32_t nm = m;
                                uint32_t opm = op->material_id & 0x3Fu;
                                switch (op->edit_type) {
                                case RKFM_EDIT_CSG_UNION: {
                                    float eff = shape_d + (1.0f - strength) * max_dim;
                                    nd = fminf(existing, eff);
                                    nm = eff < existing ? opm : ((nd < 0.0f && m == 0u) ? opm : m);
                                    break;
                                }
                                case RKFM_EDIT_CSG_SUBTRACT: {
                                    float eff = shape_d + (1.0f - strength) * max_dim;
                                    nd = fmaxf(existing, -eff);
                                    break;
                                }
                                case RKFM_EDIT_CSG_INTERSECT: {
                                    float eff = shape_d - (1.0f - strength) * max_dim;
                                    nd = fmaxf(existing, eff);
                                    break;
                                }
                                case RKFM_EDIT_SMOOTH_UNION: {
                                    float kk = fmaxf(op->blend_k, 0.001f);
                                    float eff = shape_d + (1.0f - strength) * max_dim;
                                    nd = e_smin(existing, eff, kk);
                                    nm = eff < existing ? opm : ((nd < 0.0f && m == 0u) ? opm : m);
                                    break;
                                }
                                case RKFM_EDIT_SMOOTH_SUBTRACT: {
                                    float kk = fmaxf(op->blend_k, 0.001f);
                                    float eff = shape_d + (1.0f - strength) * max_dim;
                                    nd = -e_smin(-existing, eff, kk);
                                    break;
                                }
                                case RKFM_EDIT_SMOOTH: nd = existing * (1.0f - strength * 0.3f); break;
                                case RKFM_EDIT_FLATTEN: nd = existing + (loc[1] - existing) * (strength * 0.5f); break;
                                case RKFM_EDIT_PAINT:
                                    if (fabsf(existing) < op->dimensions[0] * 2.0f) {
                                        mat[vi] = (uint8_t)opm;
                                        any = 1;
                                    }
                                    continue;
                                default: continue;
                                }
                                dist[vi] = orc_f32_to_f16(nd);
                                mat[vi] = (uint8_t)nm;
                                if (fabsf(nd - existing) > thr) any = 1;
                            }
                    if (any) modified[slot] = 1;
                }
    }
}

void orc_sparse_tiles(const orc_sparse_scene *s, const rkf_camera_uniforms *cam, const rkf_scene_uniforms *su,
                      uint32_t w, uint32_t h, int prec, double *out) {
    uint64_t n = (uint64_t)((w + 7u) / 8u) * ((h + 7u) / 8u) * 2u;
    if (prec == 64) {
        sp_tiles_f64(s, cam, (double)su->max_distance, w, h, out, NULL);
        return;
    }
    float *t = (float *)malloc(n * sizeof(float));
    sp_tiles_f32(s, cam, su->max_distance, w, h, t, NULL);
    for (uint64_t i = 0; i < n; i++) out[i] = (double)t[i];
    free(t);
}

void orc_sparse_tile_counts(const orc_sparse_scene *s, const rkf_camera_uniforms *cam,
                            const rkf_scene_uniforms *su, uint32_t w, uint32_t h, uint32_t *counts) {
    float *t = (float *)malloc((uint64_t)((w + 7u) / 8u) * ((h + 7u) / 8u) * 2u * sizeof(float));
    sp_tiles_f32(s, cam, su->max_distance, w, h, t, counts);
    free(t);
}

static int tile_cull(const rkfm_scene_ext *ext) { return ext && (ext->march_flags & RKFM_MARCH_TILE_CULL); }

void orc_sparse_trace(const orc_sparse_scene *s, const rkf_camera_uniforms *cam, const rkf_scene_uniforms *su,
                      const rkfm_scene_ext *ext, uint32_t w, uint32_t h, int prec, uint64_t p0, uint64_t p1,
                      orc_ray_result *out) {
    uint64_t nt = (uint64_t)((w + 7u) / 8u) * ((h + 7u) / 8u) * 2u;
    float *t32 = NULL;
    double *t64 = NULL;
    if (tile_cull(ext)) {
        if (prec == 64) {
            t64 = (double *)malloc(nt * sizeof(double));
            sp_tiles_f64(s, cam, (double)su->max_distance, w, h, t64, NULL);
        } else {
            t32 = (float *)malloc(nt * sizeof(float));
            sp_tiles_f32(s, cam, su->max_distance, w, h, t32, NULL);
        }
    }
    for (uint64_t p = p0; p < p1; p++) {
        if (prec == 64) sp_pixel_f64(s, cam, su, ext, w, h, t64, p, out + (p - p0));
        else sp_pixel_f32(s, cam, su, ext, w, h, t32, p, out + (p - p0));
    }
    free(t32);
    free(t64);
}

int orc_sparse_trace_parallel(const orc_sparse_scene *s, const rkf_camera_uniforms *cam,
                              const rkf_scene_uniforms *su, const rkfm_scene_ext *ext, uint32_t w, uint32_t h,
                              uint64_t p0, uint64_t p1, orc_ray_result *out, int threads) {
    int used = 1;
    float *t32 = NULL;
    if (tile_cull(ext)) {
        t32 = (float *)malloc((uint64_t)((w + 7u) / 8u) * ((h + 7u) / 8u) * 2u * sizeof(float));
        sp_tiles_f32(s, cam, su->max_distance, w, h, t32, NULL);
    }
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel
    {
#pragma omp single
        used = omp_get_num_threads();
#pragma omp for schedule(dynamic, 64)
        for (int64_t p = (int64_t)p0; p < (int64_t)p1; p++)
            sp_pixel_f32(s, cam, su, ext, w, h, t32, (uint64_t)p, out + (p - (int64_t)p0));
    }
#else
    (void)threads;
    for (uint64_t p = p0; p < p1; p++) sp_pixel_f32(s, cam, su, ext, w, h, t32, p, out + (p - p0));
#endif
    free(t32);
    return used;
}

void orc_sparse_query(const orc_sparse_scene *s, const float *xyz, uint64_t n, int prec, rkfm_point_result *out) {
    for (uint64_t i = 0; i < n; i++) {
        if (prec == 64) sp_query_f64(s, xyz + 3 * i, out + i);
        else sp_query_f32(s, xyz + 3 * i, out + i);
    }
}

/* soft_shadow (shade_common_shading.wgsl:9-33) from given origins toward L. */
void orc_sparse_shadow(const orc_sparse_scene *s, const float *origins, uint64_t n, const float L[3], float maxd,
                       float k, int prec, float *vis) {
    for (uint64_t i = 0; i < n; i++) {
        uint32_t steps;
        if (prec == 64) {
            sctx_f64 c;
            sp_init_f64(&c, s);
            vis[i] = (float)sp_shadow_f64(&c, mk_f64(origins[3 * i], origins[3 * i + 1], origins[3 * i + 2]),
                                          mk_f64(L[0], L[1], L[2]), maxd, k, &steps);
        } else {
            sctx_f32 c;
            sp_init_f32(&c, s);
            vis[i] = sp_shadow_f32(&c, mk_f32(origins[3 * i], origins[3 * i + 1], origins[3 * i + 2]),
                                   mk_f32(L[0], L[1], L[2]), maxd, k, &steps);
        }
    }
}

double orc_sparse_sample(const orc_sparse_scene *s, const float p[3], int prec) {
    if (prec == 64) {
        sctx_f64 c;
        sp_init_f64(&c, s);
        return sp_sample_f64(&c, mk_f64(p[0], p[1], p[2]), 0, 0);
    }
    sctx_f32 c;
    sp_init_f32(&c, s);
    return (double)sp_sample_f32(&c, mk_f32(p[0], p[1], p[2]), 0, 0);
}

/* G-buffer packing (ray_march.wgsl:1083-1111). Miss: (0,0,0,MAX_FLOAT), 0, 0, 0. */
void orc_pack_gbuffer(const orc_ray_result *r, uint64_t n, float *position, uint16_t *normal, uint32_t *material,
                      float *motion) {
    for (uint64_t i = 0; i < n; i++) {
        const orc_ray_result *x = r + i;
        if (x->hit) {
            position[4 * i + 0] = (float)x->pos[0];
            position[4 * i + 1] = (float)x->pos[1];
            position[4 * i + 2] = (float)x->pos[2];
            position[4 * i + 3] = (float)x->t;
            normal[4 * i + 0] = orc_f32_to_f16((float)x->normal[0]);
            normal[4 * i + 1] = orc_f32_to_f16((float)x->normal[1]);
            normal[4 * i + 2] = orc_f32_to_f16((float)x->normal[2]);
            normal[4 * i + 3] = 0;
            material[i] = x->packed_material;
            motion[4 * i + 0] = (float)x->motion[0];
            motion[4 * i + 1] = (float)x->motion[1];
            motion[4 * i + 2] = 1.0f;
            motion[4 * i + 3] = (float)x->sdf_at_hit;
        } else {
            position[4 * i + 0] = position[4 * i + 1] = position[4 * i + 2] = 0.0f;
            position[4 * i + 3] = 3.402823e+38f;
            normal[4 * i + 0] = normal[4 * i + 1] = normal[4 * i + 2] = normal[4 * i + 3] = 0;
            material[i] = 0;
            motion[4 * i + 0] = motion[4 * i + 1] = motion[4 * i + 2] = motion[4 * i + 3] = 0.0f;
        }
    }
}

/* ---- producers ---- */

/* coarse_field.rs:76-79 aabb_unsigned_distance, :240-266 populate. */
void orc_coarse_populate(const float origin[3], const uint32_t dims[3], float vs, const rkf_gpu_object *objs,
                         uint32_t n, uint16_t *cells) {
    uint32_t dmax = dims[0] > dims[1] ? dims[0] : dims[1];
    dmax = dmax > dims[2] ? dmax : dims[2];
    for (uint32_t z = 0; z < dims[2]; z++)
        for (uint32_t y = 0; y < dims[1]; y++)
            for (uint32_t x = 0; x < dims[0]; x++) {
                float wp[3] = {origin[0] + ((float)x + 0.5f) * vs, origin[1] + ((float)y + 0.5f) * vs,
                               origin[2] + ((float)z + 0.5f) * vs};
                float md = vs * (float)dmax;
                for (uint32_t i = 0; i < n; i++) {
                    float d[3];
                    for (int a = 0; a < 3; a++)
                        d[a] = fmaxf(fmaxf(objs[i].aabb_min[a] - wp[a], wp[a] - objs[i].aabb_max[a]), 0.0f);
                    float l = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                    md = fminf(md, l);
                }
                cells[(uint64_t)x + (uint64_t)y * dims[0] + (uint64_t)z * dims[0] * dims[1]] = orc_f32_to_f16(md);
            }
}

/* tile_object_cull.wgsl:90-114 project_to_pixel */
static void project(const rkf_camera_uniforms *c, const float cr[3], float out[3]) {
    float fx = c->forward[0], fy = c->forward[1], fz = c->forward[2];
    float depth = cr[0] * fx + cr[1] * fy + cr[2] * fz;
    if (depth <= 0.0f) { out[0] = -1.0f; out[1] = -1.0f; out[2] = depth; return; }
    const float *r = c->right, *u = c->up;
    float rsq = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    float usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    float nx = (cr[0] * r[0] + cr[1] * r[1] + cr[2] * r[2]) / (depth * rsq);
    float ny = (cr[0] * u[0] + cr[1] * u[1] + cr[2] * u[2]) / (depth * usq);
    out[0] = (nx * 0.5f + 0.5f) * c->resolution[0];
    out[1] = (0.5f - ny * 0.5f) * c->resolution[1];
    out[2] = depth;
}

/* tile_object_cull.wgsl:117-204 */
void orc_tile_cull(const rkf_camera_uniforms *cam, const rkf_scene_uniforms *su, const rkf_gpu_object *objs,
                   uint32_t w, uint32_t h, uint32_t *idx, uint32_t *counts) {
    uint32_t ntx = (w + 15u) / 16u, nty = (h + 15u) / 16u;
    for (uint32_t ty = 0; ty < nty; ty++)
        for (uint32_t tx = 0; tx < ntx; tx++) {
            float tmin[2] = {(float)(tx * 16u), (float)(ty * 16u)};
            float tmax[2] = {tmin[0] + 16.0f, tmin[1] + 16.0f};
            uint32_t cnt = 0;
            uint32_t tid = ty * ntx + tx;
            for (uint32_t i = 0; i < su->num_objects && cnt < 32u; i++) {
                const rkf_gpu_object *o = &objs[i];
                if (o->sdf_type == RKF_SDF_TYPE_NONE) continue;
                float smin[2] = {cam->resolution[0], cam->resolution[1]}, smax[2] = {0.0f, 0.0f};
                int front = 0, behind = 0;
                for (uint32_t c = 0; c < 8; c++) {
                    float corner[3] = {(c & 1u) ? o->aabb_max[0] : o->aabb_min[0],
                                       (c & 2u) ? o->aabb_max[1] : o->aabb_min[1],
                                       (c & 4u) ? o->aabb_max[2] : o->aabb_min[2]};
                    float cr[3] = {corner[0] - cam->position[0], corner[1] - cam->position[1],
                                   corner[2] - cam->position[2]};
                    float pr[3];
                    project(cam, cr, pr);
                    if (pr[2] > 0.0f) {
                        front = 1;
                        smin[0] = fminf(smin[0], pr[0]);
                        smin[1] = fminf(smin[1], pr[1]);
                        smax[0] = fmaxf(smax[0], pr[0]);
                        smax[1] = fmaxf(smax[1], pr[1]);
                    } else {
                        behind = 1;
                    }
                }
                int ov = 0;
                if (behind && front) ov = 1;
                else if (!behind && front)
                    ov = smax[0] >= tmin[0] && smin[0] < tmax[0] && smax[1] >= tmin[1] && smin[1] < tmax[1];
                if (ov) idx[tid * 32u + cnt++] = i;
            }
            counts[tid] = cnt;
        }
}

/*
 * core.c — oracle restatement of rkf-core: binary16, VoxelSample, index math,
 * pool slot order, CPU brick sampling, analytic SDFs, voxelize_sdf, and the
 * .rkf v2 / LZ4 block reader. Test infrastructure only (see oracle.h).
 * Build: -O2 -ffp-contract=off (no FMA contraction: Rust f32 never contracts).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* IEEE 754 binary32 -> binary16, round to nearest even; overflow -> inf,
 * NaN stays NaN (quiet). Same results as the `half` crate's f16::from_f32
 * used by VoxelSample::new (voxel.rs:33). */
uint16_t orc_f32_to_f16(float f) {
    uint32_t x = f2u(f);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t exp = (x >> 23) & 0xFFu;
    uint32_t man = x & 0x7FFFFFu;
    if (exp == 0xFFu) {
        if (man == 0) return (uint16_t)(sign | 0x7C00u);
        return (uint16_t)(sign | 0x7E00u | (man >> 13));
    }
    int e = (int)exp - 127 + 15;
    if (e >= 0x1F) return (uint16_t)(sign | 0x7C00u);
    if (e <= 0) {
        if (e < -10) return (uint16_t)sign;
        man |= 0x800000u;
        uint32_t shift = (uint32_t)(14 - e);
        uint32_t hm = man >> shift;
        uint32_t rb = 1u << (shift - 1);
        if ((man & rb) && (man & (3u * rb - 1u))) hm++;
        return (uint16_t)(sign | hm);
    }
    uint32_t h = sign | ((uint32_t)e << 10) | (man >> 13);
    uint32_t rb = 0x1000u;
    if ((man & rb) && (man & (3u * rb - 1u))) h++; /* carries into exponent */
    return (uint16_t)h;
}

/* binary16 -> binary32, exact (WGSL unpack2x16float, ray_march.wgsl:171-173). */
float orc_f16_to_f32(uint16_t h) {
    uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    uint32_t exp = ((uint32_t)h >> 10) & 0x1Fu;
    uint32_t man = (uint32_t)h & 0x3FFu;
    if (exp == 0x1Fu) return u2f(sign | 0x7F800000u | (man << 13));
    if (exp == 0) {
        if (man == 0) return u2f(sign);
        int e = -1;
        do { man <<= 1; e++; } while (!(man & 0x400u));
        man &= 0x3FFu;
        return u2f(sign | ((uint32_t)(127 - 15 - e) << 23) | (man << 13));
    }
    return u2f(sign | ((exp + 112u) << 23) | (man << 13));
}

void orc_f32_to_f16_array(const float *in, uint16_t *out, uint64_t n) {
    for (uint64_t i = 0; i < n; i++) out[i] = orc_f32_to_f16(in[i]);
}
void orc_f16_to_f32_array(const uint16_t *in, float *out, uint64_t n) {
    for (uint64_t i = 0; i < n; i++) out[i] = orc_f16_to_f32(in[i]);
}

/* ---- VoxelSample (voxel.rs:32-90) ---- */
rkf_voxel_sample orc_voxel_new(float d, uint16_t mat, uint8_t blend) {
    rkf_voxel_sample v;
    v.word0 = (uint32_t)orc_f32_to_f16(d) | (((uint32_t)mat & 0x3Fu) << 16);
    v.word1 = (uint32_t)blend << 24;
    return v;
}
rkf_voxel_sample orc_voxel_new_blended(float d, uint16_t mat, uint8_t sec, uint8_t blend) {
    rkf_voxel_sample v;
    v.word0 = (uint32_t)orc_f32_to_f16(d) | (((uint32_t)mat & 0x3Fu) << 16) |
              (((uint32_t)sec & 0x3Fu) << 22);
    v.word1 = (uint32_t)blend << 24;
    return v;
}
float orc_voxel_distance(rkf_voxel_sample v) { return orc_f16_to_f32((uint16_t)(v.word0 & 0xFFFFu)); }
uint32_t orc_voxel_material(rkf_voxel_sample v) { return (v.word0 >> 16) & 0x3Fu; }
uint32_t orc_voxel_secondary(rkf_voxel_sample v) { return (v.word0 >> 22) & 0x3Fu; }
uint32_t orc_voxel_blend(rkf_voxel_sample v) { return v.word1 >> 24; }

/* ---- brick.rs:28-42 / brick_map.rs:80-85 ---- */
int64_t orc_brick_index(uint32_t x, uint32_t y, uint32_t z) {
    if (x >= 8 || y >= 8 || z >= 8) return -1;
    return (int64_t)(x + y * 8u + z * 64u);
}
int64_t orc_map_flatten(const uint32_t d[3], uint32_t bx, uint32_t by, uint32_t bz) {
    if (bx >= d[0] || by >= d[1] || bz >= d[2]) return -1;
    return (int64_t)(bx + by * d[0] + bz * d[0] * d[1]);
}

/* ---- brick_pool.rs:37-47 (free list reversed so pop yields 0,1,2...) ---- */
int orc_pool_init(orc_pool *p, uint32_t cap) {
    p->capacity = cap;
    p->free_list = (uint32_t *)malloc(sizeof(uint32_t) * (cap ? cap : 1));
    if (!p->free_list) return -1;
    for (uint32_t i = 0; i < cap; i++) p->free_list[i] = cap - 1 - i;
    p->n_free = cap;
    return 0;
}
void orc_pool_free(orc_pool *p) { free(p->free_list); p->free_list = NULL; p->n_free = 0; }
int64_t orc_pool_allocate(orc_pool *p) {
    if (p->n_free == 0) return -1;
    return (int64_t)p->free_list[--p->n_free];
}
/* brick_pool.rs:119-129: all-or-nothing. */
int orc_pool_allocate_range(orc_pool *p, uint32_t count, uint32_t *slots) {
    if (p->n_free < count) return -1;
    for (uint32_t i = 0; i < count; i++) slots[i] = p->free_list[--p->n_free];
    return 0;
}

/* ---- sampling.rs:23-66 ---- */
static float clampf_(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }

float orc_sample_brick_trilinear(const rkf_voxel_sample *b, const float lp[3]) {
    const float dim = 8.0f;
    float fx = clampf_(lp[0] * dim - 0.5f, 0.0f, dim - 1.0001f);
    float fy = clampf_(lp[1] * dim - 0.5f, 0.0f, dim - 1.0001f);
    float fz = clampf_(lp[2] * dim - 0.5f, 0.0f, dim - 1.0001f);
    uint32_t ix = (uint32_t)floorf(fx), iy = (uint32_t)floorf(fy), iz = (uint32_t)floorf(fz);
    uint32_t ix1 = ix + 1 < 7 ? ix + 1 : 7, iy1 = iy + 1 < 7 ? iy + 1 : 7, iz1 = iz + 1 < 7 ? iz + 1 : 7;
    float tx = fx - (float)ix, ty = fy - (float)iy, tz = fz - (float)iz;
#define D(x, y, z) orc_voxel_distance(b[(x) + (y) * 8 + (z) * 64])
    float c000 = D(ix, iy, iz), c100 = D(ix1, iy, iz), c010 = D(ix, iy1, iz), c110 = D(ix1, iy1, iz);
    float c001 = D(ix, iy, iz1), c101 = D(ix1, iy, iz1), c011 = D(ix, iy1, iz1), c111 = D(ix1, iy1, iz1);
#undef D
    float c00 = c000 * (1.0f - tx) + c100 * tx;
    float c10 = c010 * (1.0f - tx) + c110 * tx;
    float c01 = c001 * (1.0f - tx) + c101 * tx;
    float c11 = c011 * (1.0f - tx) + c111 * tx;
    float c0 = c00 * (1.0f - ty) + c10 * ty;
    float c1 = c01 * (1.0f - ty) + c11 * ty;
    return c0 * (1.0f - tz) + c1 * tz;
}

/* sampling.rs:73-81: f32::round (half away from zero), `as u32` saturates. */
static uint32_t round_sat_u32(float x) {
    float r = roundf(x);
    if (!(r > 0.0f)) return 0;
    if (r >= 4294967295.0f) return 0xFFFFFFFFu;
    return (uint32_t)r;
}
uint32_t orc_sample_brick_nearest_material(const rkf_voxel_sample *b, const float lp[3]) {
    const float dim = 8.0f;
    uint32_t ix = round_sat_u32(lp[0] * dim - 0.5f), iy = round_sat_u32(lp[1] * dim - 0.5f),
             iz = round_sat_u32(lp[2] * dim - 0.5f);
    if (ix > 7) ix = 7;
    if (iy > 7) iy = 7;
    if (iz > 7) iz = 7;
    return orc_voxel_material(b[ix + iy * 8 + iz * 64]);
}

/* ---- sdf.rs:12-51 ---- */
static float len3(float x, float y, float z) { return sqrtf(x * x + y * y + z * z); }
float orc_sphere_sdf(const float c[3], float r, const float p[3]) {
    return len3(p[0] - c[0], p[1] - c[1], p[2] - c[2]) - r;
}
float orc_box_sdf(const float h[3], const float p[3]) {
    float qx = fabsf(p[0]) - h[0], qy = fabsf(p[1]) - h[1], qz = fabsf(p[2]) - h[2];
    float outside = len3(fmaxf(qx, 0.0f), fmaxf(qy, 0.0f), fmaxf(qz, 0.0f));
    float inside = fminf(fmaxf(qx, fmaxf(qy, qz)), 0.0f);
    return outside + inside;
}
float orc_capsule_sdf(const float a[3], const float b[3], float r, const float p[3]) {
    float pa[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
    float ba[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    float h = (pa[0] * ba[0] + pa[1] * ba[1] + pa[2] * ba[2]) / (ba[0] * ba[0] + ba[1] * ba[1] + ba[2] * ba[2]);
    h = clampf_(h, 0.0f, 1.0f);
    return len3(pa[0] - ba[0] * h, pa[1] - ba[1] * h, pa[2] - ba[2] * h) - r;
}
float orc_smin(float a, float b, float k) {
    float h = fmaxf(k - fabsf(a - b), 0.0f) / k;
    return fminf(a, b) - h * h * k * 0.25f;
}
float orc_smooth_min_cross(float a, float b, float k) {
    float h = fmaxf(k - fabsf(a - b), 0.0f) / fmaxf(k, 1e-6f);
    return fminf(a, b) - h * h * k * 0.25f;
}

/* fBm terrain (prim type 3, defined in include/rkf_march.h): value noise
 * on a splitmix64-hashed integer lattice, smoothstep interpolation, 6
 * octaves; d = p2 * ((y - cy) - p0 * h) as terrain.rs:229-246 (y - height). */
static uint64_t terrain_mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static float terrain_lattice(uint32_t seed, int32_t i, int32_t j) {
    uint64_t key = (uint64_t)(uint32_t)i | ((uint64_t)(uint32_t)j << 32);
    uint64_t h = terrain_mix64(key ^ ((uint64_t)seed * 0x9E3779B97F4A7C15ull));
    return (float)(uint32_t)(h >> 40) * (1.0f / 8388608.0f) - 1.0f;
}
static float terrain_vnoise(uint32_t seed, float u, float v) {
    float fu = floorf(u), fv = floorf(v);
    int32_t iu = (int32_t)fu, iv = (int32_t)fv;
    float tu = u - fu, tv = v - fv;
    float su = tu * tu * (3.0f - 2.0f * tu), sv = tv * tv * (3.0f - 2.0f * tv);
    float c00 = terrain_lattice(seed, iu, iv), c10 = terrain_lattice(seed, iu + 1, iv);
    float c01 = terrain_lattice(seed, iu, iv + 1), c11 = terrain_lattice(seed, iu + 1, iv + 1);
    float a = c00 + (c10 - c00) * su;
    float b = c01 + (c11 - c01) * su;
    return a + (b - a) * sv;
}
float orc_terrain_height(const rkfm_prim *pr, float x, float z) {
    float u = x - pr->centre[0], w = z - pr->centre[2];
    float h = 0.0f, amp = 1.0f, freq = pr->params[1];
    for (uint32_t i = 0; i < 6u; i++) {
        h = h + amp * terrain_vnoise(pr->seed + i, u * freq, w * freq);
        amp = amp * 0.5f;
        freq = freq * 2.0f;
    }
    return pr->params[0] * h;
}

/* Scene-builder primitive: translate, rotate (row-major world->local), then
 * evaluate_primitive (voxelize_object.rs:439-474) for sphere/box/capsule. */
float orc_prim_eval(const rkfm_prim *pr, const float p[3]) {
    if (pr->type == 3u) return pr->params[2] * ((p[1] - pr->centre[1]) - orc_terrain_height(pr, p[0], p[2]));
    float d0 = p[0] - pr->centre[0], d1 = p[1] - pr->centre[1], d2 = p[2] - pr->centre[2];
    const float *R = pr->rot;
    float l[3];
    l[0] = R[0] * d0 + R[1] * d1 + R[2] * d2;
    l[1] = R[3] * d0 + R[4] * d1 + R[5] * d2;
    l[2] = R[6] * d0 + R[7] * d1 + R[8] * d2;
    switch (pr->type) {
    case 0: return len3(l[0], l[1], l[2]) - pr->params[0];
    case 1: return orc_box_sdf(pr->params, l);
    case 2: {
        float a[3] = {0.0f, -pr->params[1], 0.0f}, b[3] = {0.0f, pr->params[1], 0.0f};
        return orc_capsule_sdf(a, b, pr->params[0], l);
    }
    default: return 3.402823e+38f;
    }
}

/* ---- voxelize_object.rs:53-182 ---- */
int orc_voxelize_sdf(orc_sdf_fn sdf, void *user, const float amin[3], const float amax[3],
                     float vs, uint32_t first_slot, uint32_t dims[3], uint32_t *map,
                     rkf_voxel_sample *bricks, uint64_t *n_bricks) {
    float bws = vs * 8.0f;
    for (int a = 0; a < 3; a++) {
        float sz = amax[a] - amin[a];
        float c = ceilf(sz / bws);
        uint32_t d = c > 0.0f ? (uint32_t)c : 0u;
        dims[a] = d < 1 ? 1 : d;
    }
    float go[3];
    for (int a = 0; a < 3; a++) go[a] = -((float)dims[a] * bws * 0.5f);
    float band = bws * 1.8f;
    uint64_t needed = 0;
    uint16_t mat;
    for (uint32_t bz = 0; bz < dims[2]; bz++)
        for (uint32_t by = 0; by < dims[1]; by++)
            for (uint32_t bx = 0; bx < dims[0]; bx++) {
                float bmin[3] = {go[0] + (float)bx * bws, go[1] + (float)by * bws, go[2] + (float)bz * bws};
                float hb = bws * 0.5f;
                float c[3] = {bmin[0] + hb, bmin[1] + hb, bmin[2] + hb};
                float d = sdf(c, user, &mat);
                int need = fabsf(d) < band;
                if (need) needed++;
                if (map) map[bx + by * dims[0] + bz * dims[0] * dims[1]] = need ? 0u : RKF_EMPTY_SLOT;
            }
    if (n_bricks) *n_bricks = needed;
    if (!map) return 0;
    uint32_t slot = first_slot;
    uint64_t local = 0;
    float half = vs * 0.5f;
    for (uint32_t bz = 0; bz < dims[2]; bz++)
        for (uint32_t by = 0; by < dims[1]; by++)
            for (uint32_t bx = 0; bx < dims[0]; bx++) {
                uint32_t *e = &map[bx + by * dims[0] + bz * dims[0] * dims[1]];
                if (*e == RKF_EMPTY_SLOT) continue;
                *e = slot++;
                float bmin[3] = {go[0] + (float)bx * bws, go[1] + (float)by * bws, go[2] + (float)bz * bws};
                rkf_voxel_sample *b = bricks ? bricks + local * 512 : NULL;
                local++;
                if (!b) continue;
                for (uint32_t vz = 0; vz < 8; vz++)
                    for (uint32_t vy = 0; vy < 8; vy++)
                        for (uint32_t vx = 0; vx < 8; vx++) {
                            float p[3] = {bmin[0] + ((float)vx * vs + half), bmin[1] + ((float)vy * vs + half),
                                          bmin[2] + ((float)vz * vs + half)};
                            float d = sdf(p, user, &mat);
                            b[vx + vy * 8 + vz * 64] = orc_voxel_new(d, mat, 0);
                        }
            }
    return 0;
}

/* ---- voxelize_to_geometry's brick classification (voxelize_object.rs:217-385) ----
 * Brick states from the brick-centre distance: deeper than the narrow band
 * -> INTERIOR; outside the band -> EMPTY; inside the band -> sampled at its
 * 512 voxel centres: no solid voxel (d <= 0) -> EMPTY, all solid -> INTERIOR,
 * else stored. The voxel words are voxelize_sdf's (orc_voxel_new of the
 * distance and material the callback returns); stored bricks are numbered
 * first_slot, first_slot + 1, ... in z-y-x order. */
int orc_voxelize_geometry(orc_sdf_fn sdf, void *user, const float amin[3], const float amax[3], float vs,
                          uint32_t first_slot, uint32_t dims[3], uint32_t *map, rkf_voxel_sample *bricks,
                          uint64_t *n_bricks) {
    float bws = vs * 8.0f;
    for (int a = 0; a < 3; a++) {
        float c = ceilf((amax[a] - amin[a]) / bws);
        uint32_t d = c > 0.0f ? (uint32_t)c : 0u;
        dims[a] = d < 1 ? 1 : d;
    }
    float go[3];
    for (int a = 0; a < 3; a++) go[a] = -((float)dims[a] * bws * 0.5f);
    float band = bws * 1.8f, half = vs * 0.5f;
    uint64_t nmap = (uint64_t)dims[0] * dims[1] * dims[2], stored = 0;
    uint32_t *state = (uint32_t *)malloc(nmap * sizeof(uint32_t));
    if (!state) return -1;
    uint16_t mat;
    uint64_t mi = 0;
    for (uint32_t bz = 0; bz < dims[2]; bz++)
        for (uint32_t by = 0; by < dims[1]; by++)
            for (uint32_t bx = 0; bx < dims[0]; bx++, mi++) {
                float bmin[3] = {go[0] + (float)bx * bws, go[1] + (float)by * bws, go[2] + (float)bz * bws};
                float hb = bws * 0.5f;
                float c[3] = {bmin[0] + hb, bmin[1] + hb, bmin[2] + hb};
                float d = sdf(c, user, &mat);
                if (d < -band) { state[mi] = RKF_INTERIOR_SLOT; continue; }
                if (!(fabsf(d) < band)) { state[mi] = RKF_EMPTY_SLOT; continue; }
                int any_solid = 0, all_solid = 1;
                for (uint32_t vz = 0; vz < 8; vz++)
                    for (uint32_t vy = 0; vy < 8; vy++)
                        for (uint32_t vx = 0; vx < 8; vx++) {
                            float p[3] = {bmin[0] + ((float)vx * vs + half), bmin[1] + ((float)vy * vs + half),
                                          bmin[2] + ((float)vz * vs + half)};
                            int solid = sdf(p, user, &mat) <= 0.0f;
                            any_solid |= solid;
                            all_solid &= solid;
                        }
                state[mi] = !any_solid ? RKF_EMPTY_SLOT : all_solid ? RKF_INTERIOR_SLOT : 0u;
                stored += state[mi] == 0u;
            }
    if (n_bricks) *n_bricks = stored;
    if (map) {
        uint32_t slot = first_slot;
        uint64_t local = 0;
        mi = 0;
        for (uint32_t bz = 0; bz < dims[2]; bz++)
            for (uint32_t by = 0; by < dims[1]; by++)
                for (uint32_t bx = 0; bx < dims[0]; bx++, mi++) {
                    if (state[mi] != 0u) { map[mi] = state[mi]; continue; }
                    map[mi] = slot++;
                    rkf_voxel_sample *b = bricks ? bricks + local * 512 : NULL;
                    local++;
                    if (!b) continue;
                    float bmin[3] = {go[0] + (float)bx * bws, go[1] + (float)by * bws, go[2] + (float)bz * bws};
                    for (uint32_t vz = 0; vz < 8; vz++)
                        for (uint32_t vy = 0; vy < 8; vy++)
                            for (uint32_t vx = 0; vx < 8; vx++) {
                                float p[3] = {bmin[0] + ((float)vx * vs + half), bmin[1] + ((float)vy * vs + half),
                                              bmin[2] + ((float)vz * vs + half)};
                                float d = sdf(p, user, &mat);
                                b[vx + vy * 8 + vz * 64] = orc_voxel_new(d, mat, 0);
                            }
                }
    }
    free(state);
    return 0;
}

/* orc_sdf_fn over one rkfm_prim (user), material = prim->material. */
float orc_prim_sdf(const float p[3], void *user, uint16_t *mat) {
    const rkfm_prim *pr = (const rkfm_prim *)user;
    *mat = (uint16_t)pr->material;
    return orc_prim_eval(pr, p);
}

/* gen_test_rkf.rs:28-56 */
static const float CROSS_C[5][3] = {{0.0f, 0.0f, 0.0f}, {0.55f, 0.0f, 0.0f}, {-0.55f, 0.0f, 0.0f},
                                    {0.0f, 0.0f, 0.55f}, {0.0f, 0.0f, -0.55f}};
static const float CROSS_R[5] = {0.35f, 0.25f, 0.25f, 0.25f, 0.25f};

float orc_test_cross_sdf(const float p[3], void *user, uint16_t *mat) {
    (void)user;
    float d = orc_sphere_sdf(CROSS_C[0], CROSS_R[0], p);
    for (int i = 1; i < 5; i++) {
        float di = orc_sphere_sdf(CROSS_C[i], CROSS_R[i], p);
        d = orc_smooth_min_cross(d, di, 0.15f);
    }
    *mat = 6;
    return d;
}

void orc_test_cross_aabb(float mn[3], float mx[3]) {
    float margin = 0.03f * 4.0f;
    for (int a = 0; a < 3; a++) { mn[a] = 3.4028235e38f; mx[a] = -3.4028235e38f; }
    for (int i = 0; i < 5; i++)
        for (int a = 0; a < 3; a++) {
            mn[a] = fminf(mn[a], CROSS_C[i][a] - (CROSS_R[i] + margin));
            mx[a] = fmaxf(mx[a], CROSS_C[i][a] + (CROSS_R[i] + margin));
        }
}

/* ---- LZ4 block format (public LZ4 spec; lz4_flex compress_prepend_size adds
 * a 4-byte little-endian uncompressed length, asset_file.rs:458,579) ---- */
int orc_lz4_decompress_block(const uint8_t *src, uint64_t n, uint8_t *dst, uint64_t cap,
                             uint64_t *out_len) {
    uint64_t ip = 0, op = 0;
    while (ip < n) {
        uint8_t tok = src[ip++];
        uint64_t lit = tok >> 4;
        if (lit == 15) {
            uint8_t b;
            do {
                if (ip >= n) return -1;
                b = src[ip++];
                lit += b;
            } while (b == 255);
        }
        if (ip + lit > n || op + lit > cap) return -2;
        memcpy(dst + op, src + ip, lit);
        ip += lit;
        op += lit;
        if (ip >= n) break; /* last sequence: literals only */
        if (ip + 2 > n) return -3;
        uint64_t off = (uint64_t)src[ip] | ((uint64_t)src[ip + 1] << 8);
        ip += 2;
        if (off == 0 || off > op) return -4;
        uint64_t ml = (tok & 15u) + 4u;
        if ((tok & 15u) == 15u) {
            uint8_t b;
            do {
                if (ip >= n) return -5;
                b = src[ip++];
                ml += b;
            } while (b == 255);
        }
        if (op + ml > cap) return -6;
        for (uint64_t i = 0; i < ml; i++) dst[op + i] = dst[op - off + i]; /* may overlap */
        op += ml;
    }
    *out_len = op;
    return 0;
}

static uint32_t rd32(const uint8_t *p) { uint32_t v; memcpy(&v, p, 4); return v; }
static uint64_t rd64(const uint8_t *p) { uint64_t v; memcpy(&v, p, 8); return v; }

/* asset_file.rs:74-150 (header 128 B, LOD entries 40 B), :534-590 (load). */
int orc_rkf_load(const uint8_t *f, uint64_t len, orc_rkf_header *h, uint32_t *map,
                 rkf_voxel_sample *bricks) {
    if (len < 128 + 40) return -1;
    if (memcmp(f, "RKF2", 4) != 0) return -2;
    memset(h, 0, sizeof(*h));
    h->version = rd32(f + 4);
    if (h->version != 2) return -3;
    h->lod_count = rd32(f + 8);
    h->material_count = rd32(f + 12);
    memcpy(h->aabb_min, f + 16, 12);
    memcpy(h->aabb_max, f + 28, 12);
    memcpy(h->material_ids, f + 60, 32);
    if (h->lod_count < 1) return -4;
    const uint8_t *e = f + 128;
    memcpy(&h->voxel_size, e, 4);
    h->brick_count = rd32(e + 4);
    h->dims[0] = rd32(e + 8);
    h->dims[1] = rd32(e + 12);
    h->dims[2] = rd32(e + 16);
    h->data_offset = rd64(e + 24);
    h->compressed_size = rd32(e + 32);
    h->uncompressed_size = rd32(e + 36);
    if (!map && !bricks) return 0;
    if (h->data_offset + h->compressed_size > len || h->compressed_size < 4) return -5;
    const uint8_t *blk = f + h->data_offset;
    uint64_t usz = rd32(blk);
    uint64_t nmap = (uint64_t)h->dims[0] * h->dims[1] * h->dims[2];
    uint64_t need = nmap * 4 + (uint64_t)h->brick_count * 512 * 8;
    if (usz < need) return -6;
    uint8_t *tmp = (uint8_t *)malloc(usz);
    if (!tmp) return -7;
    uint64_t got = 0;
    int rc = orc_lz4_decompress_block(blk + 4, h->compressed_size - 4, tmp, usz, &got);
    if (rc != 0 || got != usz) { free(tmp); return -8; }
    if (map) memcpy(map, tmp, nmap * 4);
    if (bricks) memcpy(bricks, tmp + nmap * 4, (uint64_t)h->brick_count * 512 * 8);
    free(tmp);
    return 0;
}

/* Scene field of the sparse builder: min over groups of the group's smin
 * chain in primitive order (sdf.rs:48-51 smin; k == 0 -> hard min). Groups
 * are identified by rkfm_prim.group; primitives of a group keep their
 * relative order in the array. Brute force over all primitives. */
float orc_scene_field(const rkfm_prim *prims, uint32_t n, float k, const float p[3]) {
    float best = INFINITY;
    for (uint32_t i = 0; i < n; i++) {
        int first = 1;
        for (uint32_t j = 0; j < i; j++)
            if (prims[j].group == prims[i].group) { first = 0; break; }
        if (!first) continue;
        float r = orc_prim_eval(&prims[i], p);
        for (uint32_t j = i + 1; j < n; j++) {
            if (prims[j].group != prims[i].group) continue;
            float dj = orc_prim_eval(&prims[j], p);
            r = k > 0.0f ? orc_smin(r, dj, k) : fminf(r, dj);
        }
        best = fminf(best, r);
    }
    return best;
}

// scene_build.cpp — host-side scene construction for the reference layouts.
//
// * voxelize: restates voxelize_sdf + populate_brick
//   (crates/rkf-core/src/voxelize_object.rs:53-182) in f32 with the Rust
//   operation order (no FMA contraction; this file is built with
//   -ffp-contract=off) and binary16 RNE encoding (VoxelSample::new voxel.rs:32-37).
// * camera: Camera::forward/right/up (camera.rs:75-93) and Camera::uniforms
//   (camera.rs:161-184); prev_vp = perspective_rh * view_matrix (camera.rs:99-110).
// * BVH: median-split tree in the GpuBvhNode layout (gpu_bvh.rs:24-41); the
//   reference builds SAH trees (bvh.rs:150-230), results are topology-invariant.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <vector>

#include "host_internal.h"

namespace {

uint16_t f32_to_f16(float f) {
    uint32_t x;
    memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u, exp = (x >> 23) & 0xFFu, man = x & 0x7FFFFFu;
    if (exp == 0xFFu) return (uint16_t)(sign | 0x7C00u | (man ? (0x200u | (man >> 13)) : 0u));
    int e = (int)exp - 112;
    if (e >= 31) return (uint16_t)(sign | 0x7C00u);
    if (e <= 0) {
        if (e < -10) return (uint16_t)sign;
        man |= 0x800000u;
        uint32_t sh = (uint32_t)(14 - e);
        uint32_t h = man >> sh, rem = man & ((1u << sh) - 1u), half = 1u << (sh - 1);
        if (rem > half || (rem == half && (h & 1u))) h++;
        return (uint16_t)(sign | h);
    }
    uint32_t h = ((uint32_t)e << 10) | (man >> 13), rem = man & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
    return (uint16_t)(sign | h);
}

rkf_voxel_sample voxel_new(float d, uint32_t mat) {
    rkf_voxel_sample v;
    v.word0 = (uint32_t)f32_to_f16(d) | ((mat & 0x3Fu) << 16);
    v.word1 = 0;
    return v;
}

// VoxelSample::new_blended (voxel.rs:43-55).
rkf_voxel_sample voxel_new_blended(float d, uint32_t mat, uint32_t sec, uint32_t blend) {
    rkf_voxel_sample v;
    v.word0 = (uint32_t)f32_to_f16(d) | ((mat & 0x3Fu) << 16) | ((sec & 0x3Fu) << 22);
    v.word1 = (blend & 0xFFu) << 24;
    return v;
}

float vlen(float x, float y, float z) { return std::sqrt(x * x + y * y + z * z); }

// Primitive SDFs of sdf.rs:12-37 for rkfm_prim types 0-2 (centre, then the
// row-major world->local rotation), the operation order of the GPU scene
// builder.
float rkf_prim_eval(const rkfm_prim &pr, float x, float y, float z) {
    const float d0 = x - pr.centre[0], d1 = y - pr.centre[1], d2 = z - pr.centre[2];
    const float *R = pr.rot;
    const float lx = R[0] * d0 + R[1] * d1 + R[2] * d2;
    const float ly = R[3] * d0 + R[4] * d1 + R[5] * d2;
    const float lz = R[6] * d0 + R[7] * d1 + R[8] * d2;
    const float *q = pr.params;
    switch (pr.type) {
    case 0: return vlen(lx, ly, lz) - q[0];
    case 1: {
        const float qx = std::fabs(lx) - q[0], qy = std::fabs(ly) - q[1], qz = std::fabs(lz) - q[2];
        const float outside = vlen(std::fmax(qx, 0.0f), std::fmax(qy, 0.0f), std::fmax(qz, 0.0f));
        return outside + std::fmin(std::fmax(qx, std::fmax(qy, qz)), 0.0f);
    }
    case 2: {
        const float hh = q[1];
        const float pax = lx - 0.0f, pay = ly - -hh, paz = lz - 0.0f;
        const float bax = 0.0f - 0.0f, bay = hh - -hh, baz = 0.0f - 0.0f;
        float h = (pax * bax + pay * bay + paz * baz) / (bax * bax + bay * bay + baz * baz);
        h = std::fmin(std::fmax(h, 0.0f), 1.0f);
        return vlen(pax - bax * h, pay - bay * h, paz - baz * h) - q[0];
    }
    default: return 3.402823e+38f;
    }
}

using SdfFn = std::function<float(float, float, float)>;
using MatFn = std::function<uint32_t(uint32_t, uint32_t, uint32_t)>;
// Secondary material | blend weight << 8 of global voxel (gx, gy, gz).
using AuxFn = std::function<uint32_t(uint32_t, uint32_t, uint32_t)>;

// The voxelize_to_geometry classification (voxelize_object.rs:217-385) over
// voxelize_sdf's grid and voxel contents: a brick whose centre lies deeper
// than the narrow band is INTERIOR; a narrow-band brick is sampled at its
// 512 voxel centres and becomes EMPTY when none is solid (d <= 0), INTERIOR
// when all are, and is stored otherwise. Stored bricks take consecutive
// slots in z-y-x order.
void voxelize_geometry(const SdfFn &sdf, const MatFn &mat, const AuxFn *aux, const float amin[3],
                       const float amax[3], float vs, uint32_t first_slot, uint32_t dims[3], uint32_t *map,
                       rkf_voxel_sample *bricks, uint64_t *n_bricks) {
    const float bws = vs * 8.0f;
    for (int a = 0; a < 3; a++) {
        float c = std::ceil((amax[a] - amin[a]) / bws);
        uint32_t d = c > 0.0f ? (uint32_t)c : 0u;
        dims[a] = std::max(d, 1u);
    }
    float go[3];
    for (int a = 0; a < 3; a++) go[a] = -((float)dims[a] * bws * 0.5f);
    const float band = bws * 1.8f, half = vs * 0.5f;
    const size_t nmap = (size_t)dims[0] * dims[1] * dims[2];
    enum : uint8_t { kEmpty, kInterior, kStored };
    std::vector<uint8_t> state(nmap, kEmpty);
    std::vector<float> vox(512);
    uint64_t count = 0;
    for (uint32_t bz = 0; bz < dims[2]; bz++)
        for (uint32_t by = 0; by < dims[1]; by++)
            for (uint32_t bx = 0; bx < dims[0]; bx++) {
                const size_t mi = bx + (size_t)by * dims[0] + (size_t)bz * dims[0] * dims[1];
                const float bx0 = go[0] + (float)bx * bws, by0 = go[1] + (float)by * bws, bz0 = go[2] + (float)bz * bws;
                const float hb = bws * 0.5f;
                const float dc = sdf(bx0 + hb, by0 + hb, bz0 + hb);
                if (dc < -band) {
                    state[mi] = kInterior;
                    continue;
                }
                if (!(std::fabs(dc) < band)) continue;
                uint32_t solid = 0;
                for (uint32_t v = 0; v < 512; v++) {
                    const float px = bx0 + ((float)(v & 7u) * vs + half);
                    const float py = by0 + ((float)((v >> 3) & 7u) * vs + half);
                    const float pz = bz0 + ((float)(v >> 6) * vs + half);
                    solid += sdf(px, py, pz) <= 0.0f;
                }
                state[mi] = solid == 0 ? kEmpty : solid == 512 ? kInterior : kStored;
                count += state[mi] == kStored;
            }
    *n_bricks = count;
    if (!map) return;
    uint32_t slot = first_slot;
    uint64_t local = 0;
    for (uint32_t bz = 0; bz < dims[2]; bz++)
        for (uint32_t by = 0; by < dims[1]; by++)
            for (uint32_t bx = 0; bx < dims[0]; bx++) {
                const size_t mi = bx + (size_t)by * dims[0] + (size_t)bz * dims[0] * dims[1];
                if (state[mi] != kStored) {
                    map[mi] = state[mi] == kInterior ? RKF_INTERIOR_SLOT : RKF_EMPTY_SLOT;
                    continue;
                }
                map[mi] = slot++;
                rkf_voxel_sample *b = bricks ? bricks + local * 512 : nullptr;
                local++;
                if (!b) continue;
                const float bx0 = go[0] + (float)bx * bws, by0 = go[1] + (float)by * bws, bz0 = go[2] + (float)bz * bws;
                for (uint32_t v = 0; v < 512; v++) {
                    const uint32_t vx = v & 7u, vy = (v >> 3) & 7u, vz = v >> 6;
                    const float d = sdf(bx0 + ((float)vx * vs + half), by0 + ((float)vy * vs + half),
                                        bz0 + ((float)vz * vs + half));
                    const uint32_t gx = bx * 8 + vx, gy = by * 8 + vy, gz = bz * 8 + vz;
                    const uint32_t m = mat(gx, gy, gz);
                    if (aux) {
                        const uint32_t a = (*aux)(gx, gy, gz);
                        b[v] = voxel_new_blended(d, m, a & 0x3Fu, a >> 8);
                    } else {
                        b[v] = voxel_new(d, m);
                    }
                }
            }
}

// voxelize_object.rs:53-157 + populate_brick :160-182.
void voxelize(const SdfFn &sdf, const MatFn &mat, const float amin[3], const float amax[3], float vs,
              uint32_t first_slot, uint32_t dims[3], uint32_t *map, rkf_voxel_sample *bricks, uint64_t *n_bricks) {
    const float bws = vs * 8.0f;
    for (int a = 0; a < 3; a++) {
        float c = std::ceil((amax[a] - amin[a]) / bws);
        uint32_t d = c > 0.0f ? (uint32_t)c : 0u;
        dims[a] = std::max(d, 1u);
    }
    float go[3];
    for (int a = 0; a < 3; a++) go[a] = -((float)dims[a] * bws * 0.5f);
    const float band = bws * 1.8f;
    std::vector<uint8_t> need((size_t)dims[0] * dims[1] * dims[2]);
    uint64_t count = 0;
    for (uint32_t bz = 0; bz < dims[2]; bz++)
        for (uint32_t by = 0; by < dims[1]; by++)
            for (uint32_t bx = 0; bx < dims[0]; bx++) {
                float hb = bws * 0.5f;
                float cx = (go[0] + (float)bx * bws) + hb;
                float cy = (go[1] + (float)by * bws) + hb;
                float cz = (go[2] + (float)bz * bws) + hb;
                bool n = std::fabs(sdf(cx, cy, cz)) < band;
                need[bx + (size_t)by * dims[0] + (size_t)bz * dims[0] * dims[1]] = n;
                count += n;
            }
    *n_bricks = count;
    if (!map) return;
    const float half = vs * 0.5f;
    uint32_t slot = first_slot;
    uint64_t local = 0;
    for (uint32_t bz = 0; bz < dims[2]; bz++)
        for (uint32_t by = 0; by < dims[1]; by++)
            for (uint32_t bx = 0; bx < dims[0]; bx++) {
                size_t mi = bx + (size_t)by * dims[0] + (size_t)bz * dims[0] * dims[1];
                if (!need[mi]) {
                    map[mi] = RKF_EMPTY_SLOT;
                    continue;
                }
                map[mi] = slot++;
                rkf_voxel_sample *b = bricks ? bricks + local * 512 : nullptr;
                local++;
                if (!b) continue;
                float bx0 = go[0] + (float)bx * bws, by0 = go[1] + (float)by * bws, bz0 = go[2] + (float)bz * bws;
                for (uint32_t vz = 0; vz < 8; vz++)
                    for (uint32_t vy = 0; vy < 8; vy++)
                        for (uint32_t vx = 0; vx < 8; vx++) {
                            float px = bx0 + ((float)vx * vs + half);
                            float py = by0 + ((float)vy * vs + half);
                            float pz = bz0 + ((float)vz * vs + half);
                            b[vx + vy * 8 + vz * 64] =
                                voxel_new(sdf(px, py, pz), mat(bx * 8 + vx, by * 8 + vy, bz * 8 + vz));
                        }
            }
}

}  // namespace

extern "C" {

int rkfm_voxelize_sphere(float radius, float margin, float vs, uint32_t object_index, uint64_t mat_seed,
                         uint32_t material, uint32_t first_slot, uint32_t dims[3], uint32_t *map_entries,
                         rkf_voxel_sample *bricks, uint64_t *n_bricks) {
    if (!dims || !n_bricks) return rkf_set_error(RKFM_ERR_INVALID_ARG, "voxelize_sphere: null argument");
    if (!(radius > 0.0f) || !(vs > 0.0f) || margin < 0.0f)
        return rkf_set_error(RKFM_ERR_INVALID_ARG, "voxelize_sphere: radius/voxel size/margin");
    float e = radius + margin;
    float amin[3] = {-e, -e, -e}, amax[3] = {e, e, e};
    SdfFn sdf = [radius](float x, float y, float z) { return vlen(x, y, z) - radius; };
    MatFn mat = [=](uint32_t gx, uint32_t gy, uint32_t gz) {
        return mat_seed ? rkf_voxel_material(mat_seed, object_index, gx, gy, gz) : material;
    };
    voxelize(sdf, mat, amin, amax, vs, first_slot, dims, map_entries, bricks, n_bricks);
    return RKFM_OK;
}

int rkfm_voxelize_prim(const rkfm_prim *prim, float margin, float vs, uint32_t flags, uint32_t object_index,
                       uint64_t mat_seed, uint32_t material, uint32_t first_slot, uint32_t dims[3],
                       uint32_t *map_entries, rkf_voxel_sample *bricks, uint64_t *n_bricks) {
    if (!prim || !dims || !n_bricks) return rkf_set_error(RKFM_ERR_INVALID_ARG, "voxelize_prim: null argument");
    if (prim->type > 2u) return rkf_set_error(RKFM_ERR_INVALID_ARG, "voxelize_prim: type %u", prim->type);
    if (!(vs > 0.0f) || margin < 0.0f || (flags & ~(RKFM_VOX_GEOMETRY_FIRST | RKFM_VOX_BLEND)))
        return rkf_set_error(RKFM_ERR_INVALID_ARG, "voxelize_prim: voxel size/margin/flags");
    const float *q = prim->params;
    const float br = prim->type == 0u ? q[0] : prim->type == 1u ? vlen(q[0], q[1], q[2]) : q[0] + q[1];
    if (!(br > 0.0f)) return rkf_set_error(RKFM_ERR_INVALID_ARG, "voxelize_prim: empty primitive");
    float amin[3], amax[3];
    for (int a = 0; a < 3; a++) {
        amin[a] = prim->centre[a] - (br + margin);
        amax[a] = prim->centre[a] + (br + margin);
    }
    const rkfm_prim pr = *prim;
    SdfFn sdf = [pr](float x, float y, float z) { return rkf_prim_eval(pr, x, y, z); };
    MatFn mat = [=](uint32_t gx, uint32_t gy, uint32_t gz) {
        return mat_seed ? rkf_voxel_material(mat_seed, object_index, gx, gy, gz) : material;
    };
    AuxFn aux = [=](uint32_t gx, uint32_t gy, uint32_t gz) {
        return rkf_voxel_blend(mat_seed, object_index, gx, gy, gz);
    };
    if (flags & RKFM_VOX_GEOMETRY_FIRST)
        voxelize_geometry(sdf, mat, (flags & RKFM_VOX_BLEND) ? &aux : nullptr, amin, amax, vs, first_slot, dims,
                          map_entries, bricks, n_bricks);
    else if (flags & RKFM_VOX_BLEND)
        return rkf_set_error(RKFM_ERR_INVALID_ARG, "voxelize_prim: BLEND needs GEOMETRY_FIRST");
    else
        voxelize(sdf, mat, amin, amax, vs, first_slot, dims, map_entries, bricks, n_bricks);
    return RKFM_OK;
}

// crates/rkf-core/examples/gen_test_rkf.rs:22-127
int rkfm_voxelize_test_cross(uint32_t dims[3], uint32_t *map_entries, rkf_voxel_sample *bricks, uint64_t *n_bricks) {
    if (!dims || !n_bricks) return rkf_set_error(RKFM_ERR_INVALID_ARG, "voxelize_test_cross: null argument");
    static const float C[5][3] = {{0.0f, 0.0f, 0.0f}, {0.55f, 0.0f, 0.0f}, {-0.55f, 0.0f, 0.0f},
                                  {0.0f, 0.0f, 0.55f}, {0.0f, 0.0f, -0.55f}};
    static const float Rr[5] = {0.35f, 0.25f, 0.25f, 0.25f, 0.25f};
    const float vs = 0.03f, k = 0.15f, margin = vs * 4.0f;
    float amin[3] = {3.4028235e38f, 3.4028235e38f, 3.4028235e38f};
    float amax[3] = {-3.4028235e38f, -3.4028235e38f, -3.4028235e38f};
    for (int i = 0; i < 5; i++)
        for (int a = 0; a < 3; a++) {
            amin[a] = std::fmin(amin[a], C[i][a] - (Rr[i] + margin));
            amax[a] = std::fmax(amax[a], C[i][a] + (Rr[i] + margin));
        }
    SdfFn sdf = [&](float x, float y, float z) {
        float d = vlen(x - C[0][0], y - C[0][1], z - C[0][2]) - Rr[0];
        for (int i = 1; i < 5; i++) {
            float di = vlen(x - C[i][0], y - C[i][1], z - C[i][2]) - Rr[i];
            float h = std::fmax(k - std::fabs(d - di), 0.0f) / std::fmax(k, 1e-6f);
            d = std::fmin(d, di) - h * h * k * 0.25f;
        }
        return d;
    };
    MatFn mat = [](uint32_t, uint32_t, uint32_t) { return 6u; };
    voxelize(sdf, mat, amin, amax, vs, 0, dims, map_entries, bricks, n_bricks);
    return RKFM_OK;
}

int rkfm_camera_look_at(const float pos[3], const float target[3], float fov_deg, uint32_t w, uint32_t h,
                        rkf_camera_uniforms *out) {
    if (!pos || !target || !out) return rkf_set_error(RKFM_ERR_INVALID_ARG, "camera_look_at: null argument");
    if (w == 0 || h == 0 || !(fov_deg > 0.0f) || !(fov_deg < 180.0f))
        return rkf_set_error(RKFM_ERR_INVALID_ARG, "camera_look_at: size/fov");
    float d[3] = {target[0] - pos[0], target[1] - pos[1], target[2] - pos[2]};
    float l = vlen(d[0], d[1], d[2]);
    if (!(l > 0.0f)) return rkf_set_error(RKFM_ERR_INVALID_ARG, "camera_look_at: target == position");
    for (float &x : d) x /= l;
    // forward = (-sin(yaw) cos(pitch), sin(pitch), -cos(yaw) cos(pitch))
    float pitch = std::asin(std::fmax(-1.0f, std::fmin(1.0f, d[1])));
    float yaw = std::atan2(-d[0], -d[2]);
    float sy = std::sin(yaw), cy = std::cos(yaw), sp = std::sin(pitch), cp = std::cos(pitch);
    float f[3] = {-sy * cp, sp, -cy * cp};
    float r[3] = {cy, 0.0f, -sy};
    // up = normalize(cross(right, forward))
    float u[3] = {r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0]};
    float ul = vlen(u[0], u[1], u[2]);
    for (float &x : u) x /= ul;
    float fov_rad = fov_deg * (3.14159265358979323846f / 180.0f);
    float hft = std::tan(fov_rad * 0.5f);
    float aspect = (float)w / (float)h;
    memset(out, 0, sizeof(*out));
    for (int a = 0; a < 3; a++) {
        out->position[a] = pos[a];
        out->forward[a] = f[a];
        out->right[a] = r[a] * hft * aspect;
        out->up[a] = u[a] * hft;
    }
    out->resolution[0] = (float)w;
    out->resolution[1] = (float)h;
    // prev_vp = P * V (camera.rs:99-110 view matrix, glam perspective_rh with
    // near 0.1, far 1000), column-major; the previous frame equals this one.
    double V[16] = {r[0], u[0], -f[0], 0, r[1], u[1], -f[1], 0, r[2], u[2], -f[2], 0,
                    -(r[0] * (double)pos[0] + r[1] * (double)pos[1] + r[2] * (double)pos[2]),
                    -(u[0] * (double)pos[0] + u[1] * (double)pos[1] + u[2] * (double)pos[2]),
                    (f[0] * (double)pos[0] + f[1] * (double)pos[1] + f[2] * (double)pos[2]), 1};
    double fy = 1.0 / std::tan(0.5 * (double)fov_rad), zn = 0.1, zf = 1000.0, rr = zf / (zn - zf);
    double Pm[16] = {fy / aspect, 0, 0, 0, 0, fy, 0, 0, 0, 0, rr, -1, 0, 0, rr * zn, 0};
    for (int c = 0; c < 4; c++)
        for (int rI = 0; rI < 4; rI++) {
            double s = 0;
            for (int k = 0; k < 4; k++) s += Pm[k * 4 + rI] * V[c * 4 + k];
            out->prev_vp[c * 4 + rI] = (float)s;
        }
    return RKFM_OK;
}

int rkfm_build_bvh(const rkf_gpu_object *objs, uint32_t n, rkf_bvh_node *nodes, uint32_t *n_nodes) {
    if (!objs || !nodes || !n_nodes || n == 0) return rkf_set_error(RKFM_ERR_INVALID_ARG, "build_bvh: argument");
    std::vector<uint32_t> idx(n);
    for (uint32_t i = 0; i < n; i++) idx[i] = i;
    uint32_t used = 0;
    std::function<uint32_t(uint32_t, uint32_t)> build = [&](uint32_t lo, uint32_t hi) -> uint32_t {
        uint32_t me = used++;
        rkf_bvh_node &nd = nodes[me];
        float mn[3] = {3.4e38f, 3.4e38f, 3.4e38f}, mx[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
        float cmn[3] = {3.4e38f, 3.4e38f, 3.4e38f}, cmx[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
        for (uint32_t i = lo; i < hi; i++) {
            const rkf_gpu_object &o = objs[idx[i]];
            for (int a = 0; a < 3; a++) {
                mn[a] = std::fmin(mn[a], o.aabb_min[a]);
                mx[a] = std::fmax(mx[a], o.aabb_max[a]);
                float c = 0.5f * (o.aabb_min[a] + o.aabb_max[a]);
                cmn[a] = std::fmin(cmn[a], c);
                cmx[a] = std::fmax(cmx[a], c);
            }
        }
        for (int a = 0; a < 3; a++) {
            nd.aabb_min[a] = mn[a];
            nd.aabb_max[a] = mx[a];
        }
        if (hi - lo == 1) {
            nd.left = RKF_BVH_INVALID;
            nd.right_or_object = idx[lo];
            return me;
        }
        int ax = 0;
        for (int a = 1; a < 3; a++)
            if (cmx[a] - cmn[a] > cmx[ax] - cmn[ax]) ax = a;
        uint32_t mid = (lo + hi) / 2;
        std::nth_element(idx.begin() + lo, idx.begin() + mid, idx.begin() + hi, [&](uint32_t a, uint32_t b) {
            float ca = objs[a].aabb_min[ax] + objs[a].aabb_max[ax], cb = objs[b].aabb_min[ax] + objs[b].aabb_max[ax];
            return ca < cb || (ca == cb && a < b);
        });
        uint32_t l = build(lo, mid);
        uint32_t r = build(mid, hi);
        nodes[me].left = l;
        nodes[me].right_or_object = r;
        return me;
    };
    build(0, n);
    *n_nodes = used;
    return RKFM_OK;
}

}  // extern "C"

// edit.hip — CSG / sculpt edits rewriting the bricks of a resident sparse
// scene.
//
// Semantics: apply_edit_cpu (crates/rkf-edit/src/cpu_apply.rs:28-188) over
// the bricks find_affected_bricks (crates/rkf-edit/src/edit_op.rs:97-140)
// selects, with the scene's global brick grid in place of one object's brick
// map: the allocated bricks (not EMPTY, not INTERIOR) in the edit's AABB
// (position +- (max dimension + blend_k)). For every voxel of such a brick:
// edit-local position through the inverse rotation, shape SDF
// (evaluate_shape, cpu_apply.rs:231-262), falloff weight (cpu_apply.rs:265-281),
// then the edit type's CSG rule on the stored distance and material, stored
// back as binary16 (RNE) and a 6-bit material. A brick is reported modified
// when some voxel's distance moved by more than 0.1 vs.
//
// Launch shape: a collect kernel walks every op's brick range (one thread
// per brick) and lists each allocated brick once (a per-slot bit dedups
// bricks shared by several ops); an apply kernel gives each listed brick a
// wavefront, eight voxels per lane, that walks the ops in order with the
// voxels in registers, so a batch of edits reads and writes each touched brick
// once; a clear kernel resets the dedup bits. No host work scales with the
// edit sizes and nothing waits for the device unless the modified list is
// asked for. Edits never allocate or free bricks, so the index and
// free-space structures stay valid.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "host_internal.h"
#include "sparse_device.h"

namespace rkf {

struct EditDev {
    rkfm_edit op;
    int32_t bmin[3], bmax[3];  // inclusive brick range (edit_op.rs:110-117)
    float max_extent;          // max dimension (cpu_apply.rs:46)
    float falloff_radius;      // max dimension + blend_k (cpu_apply.rs:73-76)
    uint32_t vol_begin, vol;   // this op's bricks in the collect grid
};

struct EditBrick {
    uint32_t slot, x, y, z;
};

RKF_DEV float e_len3(float x, float y, float z) { return __builtin_sqrtf(x * x + y * y + z * z); }

// glam Quat * Vec3 (scalar form): v (w^2 - b.b) + b (2 v.b) + (b x v)(2 w).
RKF_DEV void quat_rotate(const float q[4], float vx, float vy, float vz, float &ox, float &oy, float &oz) {
    float bx = q[0], by = q[1], bz = q[2], w = q[3];
    float b2 = bx * bx + by * by + bz * bz;
    float s = w * w - b2;
    float d2 = (vx * bx + vy * by + vz * bz) * 2.0f;
    float cx = by * vz - bz * vy, cy = bz * vx - bx * vz, cz = bx * vy - by * vx;
    float w2 = w * 2.0f;
    ox = (vx * s + bx * d2) + cx * w2;
    oy = (vy * s + by * d2) + cy * w2;
    oz = (vz * s + bz * d2) + cz * w2;
}

// evaluate_shape (cpu_apply.rs:231-262)
RKF_DEV float edit_shape(uint32_t shape, const float *dims, float px, float py, float pz) {
    switch (shape) {
    case RKFM_SHAPE_SPHERE: return e_len3(px, py, pz) - dims[0];
    case RKFM_SHAPE_BOX: {
        float qx = fabsf(px) - dims[0], qy = fabsf(py) - dims[1], qz = fabsf(pz) - dims[2];
        float outside = e_len3(fmaxf(qx, 0.0f), fmaxf(qy, 0.0f), fmaxf(qz, 0.0f));
        float inside = fminf(fmaxf(qx, fmaxf(qy, qz)), 0.0f);
        return outside + inside;
    }
    case RKFM_SHAPE_CAPSULE: {
        float ax = 0.0f, ay = -dims[1], az = 0.0f, bx = 0.0f, by = dims[1], bz = 0.0f;
        float pax = px - ax, pay = py - ay, paz = pz - az;
        float bax = bx - ax, bay = by - ay, baz = bz - az;
        float h = (pax * bax + pay * bay + paz * baz) / (bax * bax + bay * bay + baz * baz);
        h = clampf(h, 0.0f, 1.0f);
        return e_len3(pax - bax * h, pay - bay * h, paz - baz * h) - dims[0];
    }
    case RKFM_SHAPE_CYLINDER: {
        float d_radial = e_len3(px, 0.0f, pz) - dims[0];
        float d_height = fabsf(py) - dims[1];
        float ox = fmaxf(d_radial, 0.0f), oy = fmaxf(d_height, 0.0f);
        float outside = __builtin_sqrtf(ox * ox + oy * oy);
        float inside = fminf(fmaxf(d_radial, d_height), 0.0f);
        return outside + inside;
    }
    case RKFM_SHAPE_TORUS: {
        float qx = e_len3(px, 0.0f, pz) - dims[0];
        return __builtin_sqrtf(qx * qx + py * py) - dims[1];
    }
    default: return py - dims[0];  // plane
    }
}

// evaluate_falloff (cpu_apply.rs:265-281)
RKF_DEV float edit_falloff(uint32_t curve, float distance, float radius) {
    if (radius <= 0.0f || distance >= radius) return 0.0f;
    float t = distance / radius;
    if (curve == RKFM_FALLOFF_LINEAR) return 1.0f - t;
    if (curve == RKFM_FALLOFF_SHARP) {
        float f = 1.0f - t;
        return f * f * f;
    }
    return 1.0f - t * t * (3.0f - 2.0f * t);
}

RKF_DEV float edit_smin(float a, float b, float k) {
    float h = fmaxf(k - fabsf(a - b), 0.0f) / k;
    return fminf(a, b) - h * h * k * 0.25f;
}

__global__ void edit_reset_kernel(uint32_t *count) { *count = 0u; }

// One thread per brick of every op's range: allocated bricks are added to the
// list once (the per-slot bit of `touched` dedups bricks shared by ops).
__global__ __launch_bounds__(256) void edit_collect_kernel(SparseDev S, const EditDev *ops, uint32_t n_ops,
                                                           uint32_t total, uint32_t *touched, EditBrick *list,
                                                           uint32_t *count) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= total) return;
    uint32_t lo = 0, hi = n_ops;
    while (hi - lo > 1u) {
        uint32_t mid = (lo + hi) >> 1;
        if (ops[mid].vol_begin <= i) lo = mid;
        else hi = mid;
    }
    const EditDev &E = ops[lo];
    const uint32_t local = i - E.vol_begin;
    const uint32_t ex = (uint32_t)(E.bmax[0] - E.bmin[0] + 1), ey = (uint32_t)(E.bmax[1] - E.bmin[1] + 1);
    const uint32_t x = (uint32_t)E.bmin[0] + local % ex, y = (uint32_t)E.bmin[1] + (local / ex) % ey;
    const uint32_t z = (uint32_t)E.bmin[2] + local / (ex * ey);
    const uint32_t slot = sp_brick_slot(S, x, y, z);
    if (slot == RKF_EMPTY_SLOT || slot == RKF_INTERIOR_SLOT) return;
    const uint32_t bit = 1u << (slot & 31u);
    const uint32_t old = atomicOr(&touched[slot >> 5], bit);
    if (old & bit) return;
    const uint32_t k = atomicAdd(count, 1u);
    EditBrick b;
    b.slot = slot;
    b.x = x;
    b.y = y;
    b.z = z;
    list[k] = b;
}

// One op of apply_edit_cpu (cpu_apply.rs:84-170) on one voxel at (px, py, pz)
// (its stored binary16 distance bits, 6-bit material and aux word); `any`
// records a distance change above thr or a paint.
RKF_DEV void edit_voxel(const EditDev &E, float px, float py, float pz, float thr, uint32_t &dbits, uint32_t &m,
                        uint32_t &ax, bool &any) {
    const rkfm_edit &op = E.op;
    float rx = px - op.position[0], ry = py - op.position[1], rz = pz - op.position[2];
    float iq[4] = {-op.rotation[0], -op.rotation[1], -op.rotation[2], op.rotation[3]};
    float lx, ly, lz;
    quat_rotate(iq, rx, ry, rz, lx, ly, lz);
    float shape_d = edit_shape(op.shape_type, op.dimensions, lx, ly, lz);
    float w = edit_falloff(op.falloff, e_len3(rx, ry, rz), E.falloff_radius);
    if (w <= 0.0f) return;
    float existing = h2f(dbits);
    float strength = op.strength * w;
    float nd = existing;
    uint32_t nm = m;
    switch (op.edit_type) {
    case RKFM_EDIT_CSG_UNION:
    case RKFM_EDIT_SMOOTH_UNION: {
        float eff = shape_d + (1.0f - strength) * E.max_extent;
        if (op.edit_type == RKFM_EDIT_CSG_UNION) {
            nd = fminf(existing, eff);
        } else {
            float kk = fmaxf(op.blend_k, 0.001f);
            nd = edit_smin(existing, eff, kk);
        }
        nm = eff < existing ? (op.material_id & 0x3Fu) : ((nd < 0.0f && m == 0u) ? (op.material_id & 0x3Fu) : m);
        break;
    }
    case RKFM_EDIT_CSG_SUBTRACT: {
        float eff = shape_d + (1.0f - strength) * E.max_extent;
        nd = fmaxf(existing, -eff);
        break;
    }
    case RKFM_EDIT_CSG_INTERSECT: {
        float eff = shape_d - (1.0f - strength) * E.max_extent;
        nd = fmaxf(existing, eff);
        break;
    }
    case RKFM_EDIT_SMOOTH_SUBTRACT: {
        float kk = fmaxf(op.blend_k, 0.001f);
        float eff = shape_d + (1.0f - strength) * E.max_extent;
        nd = -edit_smin(-existing, eff, kk);
        break;
    }
    case RKFM_EDIT_SMOOTH: nd = existing * (1.0f - strength * 0.3f); break;
    case RKFM_EDIT_FLATTEN: nd = existing + (ly - existing) * (strength * 0.5f); break;
    case RKFM_EDIT_PAINT:
        if (fabsf(existing) < op.dimensions[0] * 2.0f) {
            m = op.material_id & 0x3Fu;
            ax = 0u;
            any = true;
            dbits = f2h(existing);  // stored again as binary16, like every written voxel
        }
        return;
    default: return;  // color paint: no geometry or material change
    }
    any |= fabsf(nd - existing) > thr;
    dbits = f2h(nd);
    m = nm;
}

// One wavefront per listed brick (grid-stride over the list, four bricks per
// 256-lane block). Lane l holds the x-row y = l & 7, z = l >> 3 (voxels
// 8l..8l+7 of the brick), read and written as one 16-byte distance, 8-byte
// material and 16-byte aux vector, so every lane has its three loads in flight
// at once; the ops are applied in order to each voxel in registers.
__global__ __launch_bounds__(256) void edit_apply_kernel(SparseDev S, const EditDev *__restrict__ ops, uint32_t n_ops,
                                                         const EditBrick *__restrict__ list,
                                                         const uint32_t *__restrict__ count, uint16_t *__restrict__ dist,
                                                         uint8_t *__restrict__ mat, uint16_t *__restrict__ aux,
                                                         uint8_t *__restrict__ out_mod) {
    const uint32_t n = *count;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t nw = gridDim.x * 4u;
    const float vs = S.vs;
    const float bws = vs * 8.0f;
    const float thr = vs * 0.1f;
    const uint32_t vy = lane & 7u, vz = lane >> 3;
    for (uint32_t e = blockIdx.x * 4u + (threadIdx.x >> 6); e < n; e += nw) {
        const EditBrick B = list[e];
        // brick_local_min = origin + b * brick_size; voxel centre = min + v * vs + vs / 2
        const float bx = S.origin[0] + (float)B.x * bws;
        const float py = (S.origin[1] + (float)B.y * bws) + (float)vy * vs + vs * 0.5f;
        const float pz = (S.origin[2] + (float)B.z * bws) + (float)vz * vs + vs * 0.5f;
        const uint64_t v0 = (uint64_t)B.slot * 512u + lane * 8u;
        const uint4 dw = *reinterpret_cast<const uint4 *>(dist + v0);
        const uint2 mw = *reinterpret_cast<const uint2 *>(mat + v0);
        const uint4 aw = aux ? *reinterpret_cast<const uint4 *>(aux + v0) : make_uint4(0u, 0u, 0u, 0u);
        const uint32_t dwv[4] = {dw.x, dw.y, dw.z, dw.w}, awv[4] = {aw.x, aw.y, aw.z, aw.w};
        uint32_t db[8], m[8], ax[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            db[j] = (dwv[j >> 1] >> (16 * (j & 1))) & 0xFFFFu;
            ax[j] = (awv[j >> 1] >> (16 * (j & 1))) & 0xFFFFu;
            m[j] = ((j < 4 ? mw.x : mw.y) >> (8 * (j & 3))) & 0xFFu;
        }
        bool any = false;
        for (uint32_t k = 0; k < n_ops; k++) {
            const EditDev &E = ops[k];
            if ((int32_t)B.x < E.bmin[0] || (int32_t)B.x > E.bmax[0] || (int32_t)B.y < E.bmin[1] ||
                (int32_t)B.y > E.bmax[1] || (int32_t)B.z < E.bmin[2] || (int32_t)B.z > E.bmax[2])
                continue;  // uniform per wavefront
#pragma unroll
            for (int j = 0; j < 8; j++)
                edit_voxel(E, bx + (float)j * vs + vs * 0.5f, py, pz, thr, db[j], m[j], ax[j], any);
        }
        uint32_t o[4], oa[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            o[i] = (db[2 * i] & 0xFFFFu) | (db[2 * i + 1] << 16);
            oa[i] = (ax[2 * i] & 0xFFFFu) | (ax[2 * i + 1] << 16);
        }
        *reinterpret_cast<uint4 *>(dist + v0) = make_uint4(o[0], o[1], o[2], o[3]);
        *reinterpret_cast<uint2 *>(mat + v0) = make_uint2(m[0] | (m[1] << 8) | (m[2] << 16) | (m[3] << 24),
                                                          m[4] | (m[5] << 8) | (m[6] << 16) | (m[7] << 24));
        if (aux) *reinterpret_cast<uint4 *>(aux + v0) = make_uint4(oa[0], oa[1], oa[2], oa[3]);
        const bool changed = __ballot(any) != 0ull;
        if (lane == 0u) out_mod[e] = changed ? 1 : 0;
    }
}

// Clears the dedup bits of the listed bricks (leaves `touched` all zero).
__global__ __launch_bounds__(256) void edit_clear_kernel(const EditBrick *list, const uint32_t *count,
                                                         uint32_t *touched) {
    const uint32_t n = *count;
    for (uint32_t e = blockIdx.x * 256u + threadIdx.x; e < n; e += gridDim.x * 256u) touched[list[e].slot >> 5] = 0u;
}

}  // namespace rkf

namespace {

bool valid_op(const rkfm_edit &op) {
    return op.edit_type <= RKFM_EDIT_COLOR_PAINT && op.shape_type <= RKFM_SHAPE_PLANE &&
           op.falloff <= RKFM_FALLOFF_SHARP;
}

}  // namespace

extern "C" int rkfm_sparse_edit(rkfm_ctx *c, rkfm_sparse_scene *s, const rkfm_edit *ops, uint32_t n_ops,
                                uint32_t *modified, uint64_t max_modified, uint64_t *n_modified, void *stream) {
    if (!c || !s || (n_ops && !ops)) return rkf_set_error(RKFM_ERR_INVALID_ARG, "sparse_edit: null argument");
    if (n_modified) *n_modified = 0;
    if (n_ops == 0 || s->info.n_bricks == 0) return RKFM_OK;
    const rkfm_sparse_info &I = s->info;
    const float bws = I.voxel_size * 8.0f;
    const int32_t nb[3] = {(int32_t)(I.cells[0] * 8u), (int32_t)(I.cells[1] * 8u), (int32_t)(I.cells[2] * 8u)};
    std::vector<rkf::EditDev> dev(n_ops);
    uint64_t total = 0;
    for (uint32_t k = 0; k < n_ops; k++) {
        const rkfm_edit &op = ops[k];
        if (!valid_op(op)) return rkf_set_error(RKFM_ERR_INVALID_ARG, "sparse_edit: op %u enum out of range", k);
        rkf::EditDev &E = dev[k];
        memset(&E, 0, sizeof(E));
        E.op = op;
        float max_dim = std::max(std::max(op.dimensions[0], op.dimensions[1]), op.dimensions[2]);
        E.max_extent = max_dim;
        E.falloff_radius = max_dim + op.blend_k;
        float extent = max_dim + op.blend_k;  // local_aabb, edit_op.rs:86-94
        uint64_t vol = 1;
        for (int a = 0; a < 3; a++) {
            float emin = op.position[a] - extent, emax = op.position[a] + extent;
            float lo = std::floor((emin - I.origin[a]) / bws);
            float hi = std::ceil((emax - I.origin[a]) / bws - 0.001f);
            if (!(std::fabs(lo) < 1e9f) || !(std::fabs(hi) < 1e9f))
                return rkf_set_error(RKFM_ERR_INVALID_ARG, "sparse_edit: op %u position / size", k);
            int32_t l = std::max((int32_t)lo, 0);
            int32_t h = std::min(std::max((int32_t)hi, 0), nb[a] - 1);
            E.bmin[a] = l;
            E.bmax[a] = h;
            vol *= l <= h ? (uint64_t)(h - l + 1) : 0u;  // empty: the edit lies off the grid
        }
        E.vol_begin = (uint32_t)total;
        E.vol = (uint32_t)vol;
        total += vol;
        if (total > (1ull << 30)) return rkf_set_error(RKFM_ERR_INVALID_ARG, "sparse_edit: edits cover too many bricks");
    }
    if (total == 0) return RKFM_OK;
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    const uint64_t ob = n_ops * sizeof(rkf::EditDev);
    int rc = c->edit_ops.reserve(ob, false);
    if (!rc) rc = c->edit_list.reserve(std::min<uint64_t>(total, I.n_bricks) * sizeof(rkf::EditBrick) + 64, false);
    if (!rc) rc = c->edit_mod.reserve(std::min<uint64_t>(total, I.n_bricks) + 64, false);
    if (!rc) rc = c->edit_count.reserve(256, false);
    if (rc) return rc;
    if (!s->edit_touched.ptr) {
        rc = s->edit_touched.reserve((I.n_bricks + 31) / 32 * 4 + 4, false);
        if (!rc) rc = rkf_hip_check(hipMemsetAsync(s->edit_touched.ptr, 0, s->edit_touched.cap, st), "edit bits clear");
        if (rc) return rc;
    }
    // Pinned staging for the op records: wait until the previous edit's copy
    // out of it has run before overwriting it.
    if (c->edit_staged) {
        rc = rkf_hip_check(hipEventSynchronize(c->edit_ev), "edit staging");
        if (rc) return rc;
    }
    if (c->edit_host_cap < ob) {
        if (c->edit_host) (void)hipHostFree(c->edit_host);
        c->edit_host = nullptr;
        c->edit_host_cap = 0;
        rc = rkf_hip_check(hipHostMalloc(&c->edit_host, ob, hipHostMallocDefault), "edit staging");
        if (rc) return rc;
        c->edit_host_cap = ob;
    }
    if (!c->edit_ev) {
        rc = rkf_hip_check(hipEventCreateWithFlags(&c->edit_ev, hipEventDisableTiming), "edit event");
        if (rc) return rc;
    }
    memcpy(c->edit_host, dev.data(), ob);
    rc = rkf_hip_check(hipMemcpyAsync(c->edit_ops.ptr, c->edit_host, ob, hipMemcpyHostToDevice, st), "edit ops upload");
    if (!rc) rc = rkf_hip_check(hipEventRecord(c->edit_ev, st), "edit event");
    if (rc) return rc;
    c->edit_staged = true;
    rkf::SparseDev D = s->dev();
    const rkf::EditDev *dops = (const rkf::EditDev *)c->edit_ops.ptr;
    uint32_t *count = (uint32_t *)c->edit_count.ptr;
    rkf::EditBrick *list = (rkf::EditBrick *)c->edit_list.ptr;
    uint32_t *touched = (uint32_t *)s->edit_touched.ptr;
    hipLaunchKernelGGL(rkf::edit_reset_kernel, dim3(1), dim3(1), 0, st, count);
    hipLaunchKernelGGL(rkf::edit_collect_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, D, dops,
                       n_ops, (uint32_t)total, touched, list, count);
    const uint64_t max_list = std::min<uint64_t>(total, I.n_bricks);
    const uint32_t apply_blocks = (uint32_t)std::min<uint64_t>((max_list + 3u) / 4u, 16384u);
    hipLaunchKernelGGL(rkf::edit_apply_kernel, dim3(apply_blocks), dim3(256), 0, st, D, dops, n_ops,
                       (const rkf::EditBrick *)list, (const uint32_t *)count, (uint16_t *)s->dist.ptr,
                       (uint8_t *)s->mat.ptr, (uint16_t *)s->aux.ptr, (uint8_t *)c->edit_mod.ptr);
    hipLaunchKernelGGL(rkf::edit_clear_kernel, dim3(256), dim3(256), 0, st, (const rkf::EditBrick *)list,
                       (const uint32_t *)count, touched);
    rc = rkf_hip_check(hipGetLastError(), "edit launch");
    if (rc || !n_modified) return rc;
    uint32_t n = 0;
    rc = rkf_hip_check(hipMemcpyAsync(&n, count, 4, hipMemcpyDeviceToHost, st), "edit count");
    if (!rc) rc = rkf_hip_check(hipStreamSynchronize(st), "edit sync");
    if (rc) return rc;
    std::vector<rkf::EditBrick> bl(n);
    std::vector<uint8_t> mod(n);
    if (n) {
        rc = rkf_hip_check(hipMemcpyAsync(bl.data(), list, (uint64_t)n * sizeof(rkf::EditBrick), hipMemcpyDeviceToHost, st),
                           "edit list");
        if (!rc) rc = rkf_hip_check(hipMemcpyAsync(mod.data(), c->edit_mod.ptr, n, hipMemcpyDeviceToHost, st), "edit flags");
        if (!rc) rc = rkf_hip_check(hipStreamSynchronize(st), "edit sync");
        if (rc) return rc;
    }
    std::vector<uint32_t> out;
    for (uint32_t i = 0; i < n; i++)
        if (mod[i]) out.push_back(bl[i].slot);
    std::sort(out.begin(), out.end());
    *n_modified = out.size();
    if (modified) memcpy(modified, out.data(), std::min<uint64_t>(out.size(), max_modified) * 4);
    return RKFM_OK;
}

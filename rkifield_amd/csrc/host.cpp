// host.cpp — context management and the C ABI of librkf_march.so.
//
// Mirrors the host side of rkf-render's ray-march pass: GpuScene uploads
// (crates/rkf-render/src/gpu_scene.rs:141-195, api/renderer/render.rs:106-153),
// the tile-cull and coarse-field producers, and RayMarchPass::dispatch
// (ray_march.rs:96-118). Scene-building helpers restate rkf-core's
// voxelize_sdf (voxelize_object.rs:53-182) and Camera::uniforms
// (camera.rs:75-93,161-184).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "host_internal.h"

static thread_local std::string g_last_error = "";

int rkf_set_error(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int rkf_hip_check(hipError_t e, const char *what) {
    if (e == hipSuccess) return RKFM_OK;
    if (e == hipErrorOutOfMemory) return rkf_set_error(RKFM_ERR_OOM, "%s: out of device memory", what);
    return rkf_set_error(RKFM_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
}

int DevBuf::reserve(uint64_t want, bool keep) {
    if (want <= cap) return RKFM_OK;
    void *np = nullptr;
    hipError_t e = hipMalloc(&np, want);
    if (e != hipSuccess) return rkf_hip_check(e, "hipMalloc");
    if (keep && ptr && size) {
        e = hipMemcpy(np, ptr, size, hipMemcpyDeviceToDevice);
        if (e != hipSuccess) {
            (void)hipFree(np);
            return rkf_hip_check(e, "hipMemcpy(grow)");
        }
    }
    if (ptr) (void)hipFree(ptr);
    ptr = np;
    cap = want;
    return RKFM_OK;
}

void DevBuf::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = size = 0;
}

extern "C" {

const char *rkfm_last_error_string(void) { return g_last_error.c_str(); }
const char *rkfm_version(void) { return "rkf-march-mi355x 0.1 (gfx950)"; }

int rkfm_create(const rkfm_config *cfg, rkfm_ctx **out) {
    if (!out) return rkf_set_error(RKFM_ERR_INVALID_ARG, "rkfm_create: out is null");
    *out = nullptr;
    int dev = cfg ? cfg->device : 0;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) return rkf_set_error(RKFM_ERR_DEVICE, "rkfm_create: no HIP device");
    if (dev < 0 || dev >= n) return rkf_set_error(RKFM_ERR_INVALID_ARG, "rkfm_create: device %d of %d", dev, n);
    e = hipSetDevice(dev);
    if (e != hipSuccess) return rkf_hip_check(e, "hipSetDevice");
    rkfm_ctx *c = new rkfm_ctx();
    c->device = dev;
    c->memory_budget = cfg ? cfg->memory_budget : 0;
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return rkf_hip_check(e, "hipStreamCreate");
    }
    (void)hipEventCreate(&c->ev0);
    (void)hipEventCreate(&c->ev1);
    int rc = rkf_noalloc_init(dev);
    if (rc) {
        (void)hipEventDestroy(c->ev0);
        (void)hipEventDestroy(c->ev1);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return rc;
    }
    c->su.num_objects = 0;
    c->su.max_steps = 128;
    c->su.max_distance = 100.0f;
    c->su.hit_threshold = 0.001f;
    memset(&c->ext, 0, sizeof(c->ext));
    *out = c;
    return RKFM_OK;
}

int rkfm_destroy(rkfm_ctx *c) {
    if (!c) return RKFM_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    rkf_sparse_release_all(c);
    c->pool.release();
    c->maps.release();
    c->objects.release();
    c->bvh.release();
    c->tile_idx.release();
    c->tile_counts.release();
    c->coarse.release();
    c->batch_items.release();
    c->batch_begins.release();
    c->batch_counter.release();
    c->batch_tiles.release();
    c->march_hits.release();
    c->march_count.release();
    c->tile_iv.release();
    c->tile_big.release();
    c->tile_big_count.release();
    c->tile_cnt.release();
    c->tile_order.release();
    c->ref_order.release();
    c->edit_ops.release();
    c->edit_list.release();
    c->edit_mod.release();
    c->edit_count.release();
    if (c->edit_host) (void)hipHostFree(c->edit_host);
    if (c->edit_ev) (void)hipEventDestroy(c->edit_ev);
    (void)hipEventDestroy(c->ev0);
    (void)hipEventDestroy(c->ev1);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return RKFM_OK;
}

int rkfm_sync(rkfm_ctx *c) {
    if (!c) return rkf_set_error(RKFM_ERR_INVALID_ARG, "rkfm_sync: null ctx");
    return rkf_hip_check(hipStreamSynchronize(c->stream), "rkfm_sync");
}

void *rkfm_stream(rkfm_ctx *c) { return c ? (void *)c->stream : nullptr; }

static int upload(rkfm_ctx *c, DevBuf &b, const void *src, uint64_t bytes, uint64_t dst_off, bool keep) {
    int rc = b.reserve(dst_off + bytes, keep);
    if (rc) return rc;
    if (bytes) {
        hipError_t e = hipMemcpyAsync((char *)b.ptr + dst_off, src, bytes, hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) return rkf_hip_check(e, "hipMemcpyAsync(upload)");
        e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return rkf_hip_check(e, "upload sync");
    }
    b.size = std::max(b.size, dst_off + bytes);
    return RKFM_OK;
}

constexpr uint64_t kMaxPoolSlots = (1ull << 23) - 1ull;

int rkfm_upload_brick_pool(rkfm_ctx *c, const rkf_voxel_sample *bricks, uint64_t n_bricks, uint64_t first_slot) {
    if (!c || (!bricks && n_bricks)) return rkf_set_error(RKFM_ERR_INVALID_ARG, "upload_brick_pool: null argument");
    // The drop-in kernel indexes voxels in 32 bits: at most 2^23 - 1 slots
    // (a 32 GiB pool; the reference's u32 indices address less).
    if (first_slot > kMaxPoolSlots || n_bricks > kMaxPoolSlots - first_slot)
        return rkf_set_error(RKFM_ERR_OUT_OF_RANGE, "upload_brick_pool: slots past %llu",
                             (unsigned long long)kMaxPoolSlots);
    const uint64_t bb = 512ull * sizeof(rkf_voxel_sample);
    int rc = upload(c, c->pool, bricks, n_bricks * bb, first_slot * bb, true);
    if (rc) return rc;
    c->n_pool_voxels = c->pool.size / sizeof(rkf_voxel_sample);
    return RKFM_OK;
}

int rkfm_upload_brick_maps(rkfm_ctx *c, const uint32_t *e, uint64_t n, uint64_t offset) {
    if (!c || (!e && n)) return rkf_set_error(RKFM_ERR_INVALID_ARG, "upload_brick_maps: null argument");
    // map entries are indexed in 32 bits (brick_map_offset is a u32)
    if (offset > 0xFFFFFFFFull || n > 0xFFFFFFFFull - offset)
        return rkf_set_error(RKFM_ERR_OUT_OF_RANGE, "upload_brick_maps: entries past 2^32 - 1");
    int rc = upload(c, c->maps, e, n * 4, offset * 4, true);
    if (rc) return rc;
    c->n_maps = c->maps.size / 4;
    return RKFM_OK;
}

int rkfm_upload_objects(rkfm_ctx *c, const rkf_gpu_object *o, uint32_t n) {
    if (!c || (!o && n)) return rkf_set_error(RKFM_ERR_INVALID_ARG, "upload_objects: null argument");
    c->objects.size = 0;
    int rc = upload(c, c->objects, o, (uint64_t)n * sizeof(rkf_gpu_object), 0, false);
    if (rc) return rc;
    c->n_objects = n;
    return RKFM_OK;
}

int rkfm_upload_bvh(rkfm_ctx *c, const rkf_bvh_node *nodes, uint32_t n) {
    if (!c || (!nodes && n)) return rkf_set_error(RKFM_ERR_INVALID_ARG, "upload_bvh: null argument");
    c->bvh.size = 0;
    int rc = upload(c, c->bvh, nodes, (uint64_t)n * sizeof(rkf_bvh_node), 0, false);
    if (rc) return rc;
    c->n_bvh = n;
    return RKFM_OK;
}

int rkfm_set_camera(rkfm_ctx *c, const rkf_camera_uniforms *cam) {
    if (!c || !cam) return rkf_set_error(RKFM_ERR_INVALID_ARG, "set_camera: null argument");
    c->cam = *cam;
    c->cam_set = true;
    return RKFM_OK;
}

int rkfm_set_scene(rkfm_ctx *c, const rkf_scene_uniforms *su, const rkfm_scene_ext *ext) {
    if (!c || !su) return rkf_set_error(RKFM_ERR_INVALID_ARG, "set_scene: null argument");
    if (ext && ext->hit_eps_mode > RKFM_EPS_RELATIVE)
        return rkf_set_error(RKFM_ERR_INVALID_ARG, "set_scene: hit_eps_mode %u", ext->hit_eps_mode);
    if (ext && ext->strategy > RKFM_STRATEGY_BRUTE)
        return rkf_set_error(RKFM_ERR_INVALID_ARG, "set_scene: strategy %u", ext->strategy);
    c->su = *su;
    if (ext) c->ext = *ext;
    else memset(&c->ext, 0, sizeof(c->ext));
    c->scene_set = true;
    return RKFM_OK;
}

// The tiled march's launch order, derived from the tile lists wherever they
// are made (tiles by descending object count; nothing else reads it).
static int ref_order(rkfm_ctx *c, uint64_t nt) {
    c->ref_order_n = 0;
    int rc = c->ref_order.reserve(nt * 4, false);
    if (rc) return rc;
    rc = rkf_hip_check(rkf_launch_ref_tile_order((const uint32_t *)c->tile_counts.ptr, (uint32_t)nt,
                                                 (uint32_t *)c->ref_order.ptr, c->stream),
                       "tile order launch");
    if (rc) return rc;
    c->ref_order_n = nt;
    return RKFM_OK;
}

int rkfm_set_tiles(rkfm_ctx *c, const uint32_t *idx, const uint32_t *counts, uint32_t tx, uint32_t ty) {
    if (!c) return rkf_set_error(RKFM_ERR_INVALID_ARG, "set_tiles: null ctx");
    c->ref_order_n = 0;
    if (tx == 0 || ty == 0) {
        c->tiles_x = c->tiles_y = 0;
        return RKFM_OK;
    }
    if (!idx || !counts) return rkf_set_error(RKFM_ERR_INVALID_ARG, "set_tiles: null lists");
    uint64_t nt = (uint64_t)tx * ty;
    c->tile_idx.size = c->tile_counts.size = 0;
    int rc = upload(c, c->tile_idx, idx, nt * RKF_TILE_MAX_OBJECTS * 4, 0, false);
    if (rc) return rc;
    rc = upload(c, c->tile_counts, counts, nt * 4, 0, false);
    if (rc) return rc;
    c->tiles_x = tx;
    c->tiles_y = ty;
    rc = ref_order(c, nt);
    if (rc) return rc;
    // a host upload: complete on return, like the lists themselves
    return rkf_hip_check(hipStreamSynchronize(c->stream), "set_tiles sync");
}

int rkfm_set_coarse_field(rkfm_ctx *c, const uint16_t *cells, const rkf_coarse_field_uniforms *u) {
    if (!c) return rkf_set_error(RKFM_ERR_INVALID_ARG, "set_coarse_field: null ctx");
    if (!cells) {
        c->coarse_set = false;
        return RKFM_OK;
    }
    if (!u) return rkf_set_error(RKFM_ERR_INVALID_ARG, "set_coarse_field: null uniforms");
    uint64_t n = (uint64_t)u->dims[0] * u->dims[1] * u->dims[2];
    if (n == 0) return rkf_set_error(RKFM_ERR_INVALID_ARG, "set_coarse_field: zero dims");
    c->coarse.size = 0;
    int rc = upload(c, c->coarse, cells, n * 2, 0, false);
    if (rc) return rc;
    c->coarse_u = *u;
    c->coarse_set = true;
    return RKFM_OK;
}

int rkfm_build_tiles(rkfm_ctx *c, uint32_t w, uint32_t h) {
    if (!c) return rkf_set_error(RKFM_ERR_INVALID_ARG, "build_tiles: null ctx");
    if (!c->cam_set || !c->scene_set) return rkf_set_error(RKFM_ERR_NOT_READY, "build_tiles: camera/scene not set");
    if (w == 0 || h == 0) return rkf_set_error(RKFM_ERR_INVALID_ARG, "build_tiles: zero size");
    if (c->su.num_objects > c->n_objects)
        return rkf_set_error(RKFM_ERR_OUT_OF_RANGE, "build_tiles: num_objects %u > uploaded %u", c->su.num_objects,
                             c->n_objects);
    uint32_t tx = (w + 15u) / 16u, ty = (h + 15u) / 16u;
    uint64_t nt = (uint64_t)tx * ty;
    c->ref_order_n = 0;
    int rc = c->tile_idx.reserve(nt * RKF_TILE_MAX_OBJECTS * 4, false);
    if (!rc) rc = c->tile_counts.reserve(nt * 4, false);
    if (rc) return rc;
    c->tile_idx.size = nt * RKF_TILE_MAX_OBJECTS * 4;
    c->tile_counts.size = nt * 4;
    rkf::TileCullParams P;
    P.objects = (const rkf_gpu_object *)c->objects.ptr;
    P.num_objects = c->su.num_objects;
    P.cam = c->cam;
    P.width = w;
    P.height = h;
    P.tiles_x = tx;
    P.tiles_y = ty;
    P.idx = (uint32_t *)c->tile_idx.ptr;
    P.counts = (uint32_t *)c->tile_counts.ptr;
    rc = rkf_hip_check(rkf_launch_tile_cull(&P, c->stream), "tile_cull launch");
    if (rc) return rc;
    c->tiles_x = tx;
    c->tiles_y = ty;
    return ref_order(c, nt);
}

int rkfm_build_coarse_field(rkfm_ctx *c, const float origin[3], const uint32_t dims[3], float vs) {
    if (!c || !origin || !dims) return rkf_set_error(RKFM_ERR_INVALID_ARG, "build_coarse_field: null argument");
    if (!c->cam_set) return rkf_set_error(RKFM_ERR_NOT_READY, "build_coarse_field: camera not set");
    if (!(vs > 0.0f)) return rkf_set_error(RKFM_ERR_INVALID_ARG, "build_coarse_field: voxel size");
    uint64_t n = (uint64_t)dims[0] * dims[1] * dims[2];
    if (n == 0) return rkf_set_error(RKFM_ERR_INVALID_ARG, "build_coarse_field: zero dims");
    int rc = c->coarse.reserve(n * 2, false);
    if (rc) return rc;
    c->coarse.size = n * 2;
    rkf::CoarseParams P;
    P.objects = (const rkf_gpu_object *)c->objects.ptr;
    P.n_objects = c->n_objects;
    for (int a = 0; a < 3; a++) {
        P.origin[a] = origin[a];
        P.dims[a] = dims[a];
    }
    P.voxel_size = vs;
    P.cells = (uint16_t *)c->coarse.ptr;
    rc = rkf_hip_check(rkf_launch_coarse(&P, c->stream), "coarse launch");
    if (rc) return rc;
    // CoarseField::upload (coarse_field.rs:279-307): camera-relative origin.
    rkf_coarse_field_uniforms u;
    memset(&u, 0, sizeof(u));
    for (int a = 0; a < 3; a++) {
        u.origin_cam_rel[a] = origin[a] - c->cam.position[a];
        u.dims[a] = dims[a];
    }
    u.voxel_size = vs;
    u.inv_voxel_size = 1.0f / vs;
    c->coarse_u = u;
    c->coarse_set = true;
    return RKFM_OK;
}

int rkfm_read_tiles(rkfm_ctx *c, uint32_t *idx, uint32_t *counts, uint64_t nt) {
    if (!c || !idx || !counts) return rkf_set_error(RKFM_ERR_INVALID_ARG, "read_tiles: null argument");
    if (nt != (uint64_t)c->tiles_x * c->tiles_y) return rkf_set_error(RKFM_ERR_OUT_OF_RANGE, "read_tiles: size");
    int rc = rkf_hip_check(hipStreamSynchronize(c->stream), "read_tiles sync");
    if (!rc) rc = rkf_hip_check(hipMemcpy(idx, c->tile_idx.ptr, nt * 128, hipMemcpyDeviceToHost), "read_tiles");
    if (!rc) rc = rkf_hip_check(hipMemcpy(counts, c->tile_counts.ptr, nt * 4, hipMemcpyDeviceToHost), "read_tiles");
    return rc;
}

int rkfm_read_coarse_field(rkfm_ctx *c, uint16_t *cells, uint64_t n, rkf_coarse_field_uniforms *u) {
    if (!c || !cells) return rkf_set_error(RKFM_ERR_INVALID_ARG, "read_coarse_field: null argument");
    if (!c->coarse_set || n * 2 != c->coarse.size)
        return rkf_set_error(RKFM_ERR_OUT_OF_RANGE, "read_coarse_field: size");
    int rc = rkf_hip_check(hipStreamSynchronize(c->stream), "read_coarse sync");
    if (!rc) rc = rkf_hip_check(hipMemcpy(cells, c->coarse.ptr, n * 2, hipMemcpyDeviceToHost), "read_coarse");
    if (!rc && u) *u = c->coarse_u;
    return rc;
}

}  // extern "C"

int rkf_fill_gbuf(const rkfm_gbuffer *out, rkf::GbufPtrs &g) {
    if (!out) return rkf_set_error(RKFM_ERR_INVALID_ARG, "dispatch: null gbuffer");
    if (!out->position || !out->normal || !out->material || !out->motion)
        return rkf_set_error(RKFM_ERR_INVALID_ARG, "dispatch: null G-buffer plane");
    if (out->width == 0 || out->height == 0) return rkf_set_error(RKFM_ERR_INVALID_ARG, "dispatch: zero size");
    g.position = (float4 *)out->position;
    g.normal = (uint2 *)out->normal;
    g.material = (uint32_t *)out->material;
    g.motion = (float4 *)out->motion;
    g.debug = (rkfm_debug_px *)out->debug;
    g.shadow = (float *)out->shadow;
    g.width = out->width;
    g.height = out->height;
    return RKFM_OK;
}

extern "C" {

int rkfm_dispatch(rkfm_ctx *c, const rkfm_gbuffer *out, void *stream) {
    if (!c) return rkf_set_error(RKFM_ERR_INVALID_ARG, "dispatch: null ctx");
    if (!c->cam_set || !c->scene_set) return rkf_set_error(RKFM_ERR_NOT_READY, "dispatch: camera/scene not set");
    rkf::RefParams P;
    memset(&P, 0, sizeof(P));
    int rc = rkf_fill_gbuf(out, P.g);
    if (rc) return rc;
    if (c->su.num_objects > c->n_objects)
        return rkf_set_error(RKFM_ERR_OUT_OF_RANGE, "dispatch: num_objects %u > uploaded %u", c->su.num_objects,
                             c->n_objects);
    uint32_t mode;
    switch (c->ext.strategy) {
    case RKFM_STRATEGY_TILED: mode = 0; break;
    case RKFM_STRATEGY_BVH: mode = 1; break;
    case RKFM_STRATEGY_BRUTE: mode = 2; break;
    default: mode = c->tiles_x ? 0u : (c->n_bvh ? 1u : 2u); break;
    }
    if (mode == 0) {
        uint32_t ntx = (out->width + 15u) / 16u, nty = (out->height + 15u) / 16u;
        if (c->tiles_x != ntx || c->tiles_y != nty)
            return rkf_set_error(RKFM_ERR_LAYOUT, "dispatch: tile grid %ux%u does not match %ux%u pixels",
                                 c->tiles_x, c->tiles_y, out->width, out->height);
    }
    // the kernel reads entry 0 of both for corners it does not use
    rc = c->pool.reserve(64, true);
    if (!rc) rc = c->maps.reserve(64, true);
    if (rc) return rc;
    P.pool = (const rkf_voxel_sample *)c->pool.ptr;
    P.n_pool_voxels = c->n_pool_voxels;
    P.maps = (const uint32_t *)c->maps.ptr;
    P.n_maps = c->n_maps;
    P.objects = (const rkf_gpu_object *)c->objects.ptr;
    P.bvh = (const rkf_bvh_node *)c->bvh.ptr;
    P.n_bvh = c->n_bvh;
    P.tile_idx = (const uint32_t *)c->tile_idx.ptr;
    P.tile_counts = (const uint32_t *)c->tile_counts.ptr;
    P.n_tiles = c->tiles_x * c->tiles_y;
    P.tile_order = mode == 0 && c->ref_order_n == P.n_tiles ? (const uint32_t *)c->ref_order.ptr : nullptr;
    P.coarse = c->coarse_set ? (const uint16_t *)c->coarse.ptr : nullptr;
    P.cu = c->coarse_u;
    P.cam = c->cam;
    P.su = c->su;
    P.eps_mode = c->ext.hit_eps_mode;
    P.mode = mode;
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    return rkf_hip_check(rkf_launch_ref_march(&P, s), "ray_march launch");
}

// ---- device memory helpers (the C ABI takes device pointers) ----
int rkfm_alloc(rkfm_ctx *c, uint64_t bytes, void **dptr) {
    if (!c || !dptr) return rkf_set_error(RKFM_ERR_INVALID_ARG, "alloc: null argument");
    *dptr = nullptr;
    if (bytes == 0) bytes = 1;
    return rkf_hip_check(hipMalloc(dptr, bytes), "hipMalloc");
}
int rkfm_free(rkfm_ctx *c, void *dptr) {
    (void)c;
    if (dptr) return rkf_hip_check(hipFree(dptr), "hipFree");
    return RKFM_OK;
}
int rkfm_memcpy_h2d(rkfm_ctx *c, void *dst, const void *src, uint64_t bytes) {
    if (!c) return rkf_set_error(RKFM_ERR_INVALID_ARG, "memcpy: null ctx");
    int rc = rkf_hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream), "memcpy h2d");
    if (!rc) rc = rkf_hip_check(hipStreamSynchronize(c->stream), "memcpy h2d sync");
    return rc;
}
int rkfm_memcpy_d2h(rkfm_ctx *c, void *dst, const void *src, uint64_t bytes) {
    if (!c) return rkf_set_error(RKFM_ERR_INVALID_ARG, "memcpy: null ctx");
    int rc = rkf_hip_check(hipStreamSynchronize(c->stream), "memcpy d2h sync");
    if (!rc) rc = rkf_hip_check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "memcpy d2h");
    return rc;
}
int rkfm_memset(rkfm_ctx *c, void *dst, int value, uint64_t bytes) {
    if (!c) return rkf_set_error(RKFM_ERR_INVALID_ARG, "memset: null ctx");
    return rkf_hip_check(hipMemsetAsync(dst, value, bytes, c->stream), "memset");
}

int rkfm_device_copy(rkfm_ctx *c, void *dst, const void *src, uint64_t bytes) {
    if (!c || !dst || !src) return rkf_set_error(RKFM_ERR_INVALID_ARG, "device_copy: null argument");
    if (bytes % 16u) return rkf_set_error(RKFM_ERR_INVALID_ARG, "device_copy: size must be a multiple of 16");
    return rkf_hip_check(rkf_launch_copy(dst, src, bytes, c->stream), "copy launch");
}

int rkfm_device_gather16(rkfm_ctx *c, void *dst, const void *src, uint64_t src_bytes, uint64_t n) {
    if (!c || !dst || !src) return rkf_set_error(RKFM_ERR_INVALID_ARG, "device_gather16: null argument");
    const uint64_t lines = src_bytes / 128u;
    if (lines == 0 || (lines & (lines - 1u)) || n > lines)
        return rkf_set_error(RKFM_ERR_INVALID_ARG, "device_gather16: %llu lines (a power of two >= n needed)",
                             (unsigned long long)lines);
    return rkf_hip_check(rkf_launch_gather16(dst, src, lines, n, c->stream), "gather16 launch");
}

// ---- timing on the context stream ----
int rkfm_timer_begin(rkfm_ctx *c) {
    if (!c) return rkf_set_error(RKFM_ERR_INVALID_ARG, "timer: null ctx");
    return rkf_hip_check(hipEventRecord(c->ev0, c->stream), "hipEventRecord");
}
int rkfm_timer_end(rkfm_ctx *c, float *ms) {
    if (!c || !ms) return rkf_set_error(RKFM_ERR_INVALID_ARG, "timer: null argument");
    int rc = rkf_hip_check(hipEventRecord(c->ev1, c->stream), "hipEventRecord");
    if (!rc) rc = rkf_hip_check(hipEventSynchronize(c->ev1), "hipEventSynchronize");
    if (!rc) rc = rkf_hip_check(hipEventElapsedTime(ms, c->ev0, c->ev1), "hipEventElapsedTime");
    return rc;
}

}  // extern "C"

"""ctypes binding of librkf_march.so and a Python mirror of rkf-render's pass API.

The classes keep the reference's names and argument meaning:

* ``GpuScene``      crates/rkf-render/src/gpu_scene.rs:52-195 (brick pool, brick
                    maps, objects, BVH, camera and scene uniforms)
* ``GBuffer``       crates/rkf-render/src/gbuffer.rs:14-58 (4 targets in the
                    reference formats, device-resident)
* ``TileObjectCullPass`` crates/rkf-render/src/tile_object_cull.rs:23-43,196
* ``CoarseField``   crates/rkf-render/src/coarse_field.rs:45-307
* ``RayMarchPass``  crates/rkf-render/src/ray_march.rs:29-118 (new / dispatch /
                    recreate_pipeline)

Errors: every C call returns a status; a non-zero status raises
``RkfMarchError`` carrying the library's message. The reference's dispatch
cannot fail from the caller's view (wgpu validation errors are logged);
here invalid inputs are reported instead of rendered.
"""
import atexit
import ctypes as C
import os
import weakref

import numpy as np

from . import layouts as L

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librkf_march.so")
_LIB = None

# Servers and contexts still open when the interpreter exits: Python does
# not promise to run __del__ for them, so they are closed here, servers
# (whose worker thread uses the context) before contexts (whose graphs are
# destroyed first), while the HIP runtime is still up.
_OPEN_SERVERS = weakref.WeakSet()
_OPEN_CONTEXTS = weakref.WeakSet()


def _close_open():
    for group in (_OPEN_SERVERS, _OPEN_CONTEXTS):
        for obj in list(group):
            try:
                obj.close()
            except Exception:
                pass


atexit.register(_close_open)


class RkfMarchError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"rkf_march error {code}: {msg}")
        self.code = code


ERR = dict(OK=0, INVALID_ARG=-1, OUT_OF_RANGE=-2, OOM=-3, DEVICE=-4, LAYOUT=-5, NOT_READY=-6, BUSY=-7, FORMAT=-8)

_VP = C.c_void_p
_U32 = C.c_uint32
_U64 = C.c_uint64
_F = C.c_float

_SIGS = {
    "rkfm_last_error_string": (C.c_char_p, []),
    "rkfm_version": (C.c_char_p, []),
    "rkfm_create": (C.c_int, [_VP, C.POINTER(_VP)]),
    "rkfm_destroy": (C.c_int, [_VP]),
    "rkfm_sync": (C.c_int, [_VP]),
    "rkfm_stream": (_VP, [_VP]),
    "rkfm_upload_brick_pool": (C.c_int, [_VP, _VP, _U64, _U64]),
    "rkfm_upload_brick_maps": (C.c_int, [_VP, _VP, _U64, _U64]),
    "rkfm_upload_objects": (C.c_int, [_VP, _VP, _U32]),
    "rkfm_upload_bvh": (C.c_int, [_VP, _VP, _U32]),
    "rkfm_set_camera": (C.c_int, [_VP, _VP]),
    "rkfm_set_scene": (C.c_int, [_VP, _VP, _VP]),
    "rkfm_set_tiles": (C.c_int, [_VP, _VP, _VP, _U32, _U32]),
    "rkfm_set_coarse_field": (C.c_int, [_VP, _VP, _VP]),
    "rkfm_build_tiles": (C.c_int, [_VP, _U32, _U32]),
    "rkfm_build_coarse_field": (C.c_int, [_VP, _VP, _VP, _F]),
    "rkfm_read_tiles": (C.c_int, [_VP, _VP, _VP, _U64]),
    "rkfm_read_coarse_field": (C.c_int, [_VP, _VP, _U64, _VP]),
    "rkfm_dispatch": (C.c_int, [_VP, _VP, _VP]),
    "rkfm_voxelize_sphere": (C.c_int, [_F, _F, _F, _U32, _U64, _U32, _U32, _VP, _VP, _VP, _VP]),
    "rkfm_voxelize_test_cross": (C.c_int, [_VP, _VP, _VP, _VP]),
    "rkfm_voxelize_prim": (C.c_int, [_VP, _F, _F, _U32, _U32, _U64, _U32, _U32, _VP, _VP, _VP, _VP]),
    "rkfm_camera_look_at": (C.c_int, [_VP, _VP, _F, _U32, _U32, _VP]),
    "rkfm_build_bvh": (C.c_int, [_VP, _U32, _VP, _VP]),
    "rkfm_sparse_build": (C.c_int, [_VP, _VP, _U32, _VP, C.POINTER(_VP)]),
    "rkfm_sparse_from_dense": (C.c_int, [_VP, _VP, _VP, _VP, _U64, _F, _VP, C.POINTER(_VP)]),
    "rkfm_sparse_from_objects": (C.c_int, [_VP, _F, _VP, _VP, C.POINTER(_VP)]),
    "rkfm_sparse_download_aux": (C.c_int, [_VP, _VP, _VP, _VP]),
    "rkfm_sparse_info_get": (C.c_int, [_VP, _VP]),
    "rkfm_sparse_destroy": (C.c_int, [_VP, _VP]),
    "rkfm_sparse_download": (C.c_int, [_VP, _VP, _VP, _VP, _VP, _VP]),
    "rkfm_sparse_download_free": (C.c_int, [_VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    "rkfm_sparse_march": (C.c_int, [_VP, _VP, _VP, _VP, _VP, _VP, _U32, _VP]),
    "rkfm_sparse_march_noalloc": (C.c_int, [_VP, _VP, _VP, _VP, _VP, _VP, _U32, _VP]),
    "rkfm_sparse_query_points": (C.c_int, [_VP, _VP, _VP, _U32, _VP, _VP]),
    "rkfm_sparse_sphere_cast": (C.c_int, [_VP, _VP, _VP, _U32, _VP, _VP]),
    "rkfm_sparse_tile_intervals": (C.c_int, [_VP, _VP, _VP, _VP, _U32, _U32, _VP, _VP]),
    "rkfm_sparse_tile_order": (C.c_int, [_VP, _VP, _VP, _VP, _U32, _U32, _VP, _VP, _VP]),
    "rkfm_sparse_edit": (C.c_int, [_VP, _VP, _VP, _U32, _VP, _U64, _VP, _VP]),
    "rkfm_sparse_contacts": (C.c_int, [_VP, _VP, _VP, _U32, _VP, _VP, _VP]),
    "rkfm_rkf_read_header": (C.c_int, [_VP, _U64, _VP, _VP, _U32]),
    "rkfm_rkf_decode_lod": (C.c_int, [_VP, _U64, _U32, _VP, _VP]),
    "rkfm_rkf_encode": (C.c_int, [_VP, _VP, _U32, _VP, _U64, _VP]),
    "rkfm_sparse_from_rkf": (C.c_int, [_VP, _VP, _U64, _U32, _VP, C.POINTER(_VP)]),
    "rkfm_alloc": (C.c_int, [_VP, _U64, C.POINTER(_VP)]),
    "rkfm_free": (C.c_int, [_VP, _VP]),
    "rkfm_memcpy_h2d": (C.c_int, [_VP, _VP, _VP, _U64]),
    "rkfm_memcpy_d2h": (C.c_int, [_VP, _VP, _VP, _U64]),
    "rkfm_memset": (C.c_int, [_VP, _VP, C.c_int, _U64]),
    "rkfm_device_copy": (C.c_int, [_VP, _VP, _VP, _U64]),
    "rkfm_device_gather16": (C.c_int, [_VP, _VP, _VP, _U64, _U64]),
    "rkfm_batch_run": (C.c_int, [_VP, _VP, _U32, _VP, _VP]),
    "rkfm_batch_autotune": (C.c_int, [_VP, _VP, _U32, _U32, _VP, C.POINTER(_F)]),
    "rkfm_batch_graph_create": (C.c_int, [_VP, _VP, _U32, _VP, C.POINTER(_VP)]),
    "rkfm_batch_graph_launch": (C.c_int, [_VP, _VP]),
    "rkfm_batch_graph_destroy": (C.c_int, [_VP]),
    "rkfm_batch_work_items": (_U64, [_VP, _U32]),
    "rkfm_host_alloc": (C.c_int, [_U64, C.POINTER(_VP)]),
    "rkfm_host_free": (C.c_int, [_VP]),
    "rkfm_server_create": (C.c_int, [_VP, _VP, _U32, _VP, C.POINTER(_VP)]),
    "rkfm_server_submit": (C.c_int, [_VP, _VP, C.POINTER(_U64)]),
    "rkfm_server_status": (C.c_int, [_VP, _U64, _VP]),
    "rkfm_server_wait": (C.c_int, [_VP, _U64, C.c_double, _VP]),
    "rkfm_server_view_planes": (C.c_int, [_VP, _U64, _VP]),
    "rkfm_server_release": (C.c_int, [_VP, _U64]),
    "rkfm_server_stats_get": (C.c_int, [_VP, _VP]),
    "rkfm_server_now": (C.c_double, [_VP]),
    "rkfm_server_destroy": (C.c_int, [_VP]),
    "rkfm_server_replay": (C.c_int, [_VP, _VP, _VP, _U32, _VP, C.POINTER(C.c_double)]),
    "rkfm_timer_begin": (C.c_int, [_VP]),
    "rkfm_timer_end": (C.c_int, [_VP, C.POINTER(_F)]),
}


def lib():
    """Load librkf_march.so; raise if it is not built (there is no fallback)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        lib_ = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(lib_, name)
            f.restype = res
            f.argtypes = args
        _LIB = lib_
    return _LIB


def exported_symbols():
    return sorted(_SIGS)


def _check(rc):
    if rc != 0:
        raise RkfMarchError(rc, lib().rkfm_last_error_string().decode())


def _p(a):
    if a is None:
        return None
    if isinstance(a, DeviceArray):
        return a.ptr
    return a.ctypes.data


class Context:
    """rkfm_ctx: owns device state on one HIP device."""

    def __init__(self, device=0, memory_budget=0):
        cfg = np.zeros(1, np.dtype([("device", "<i4"), ("flags", "<u4"), ("memory_budget", "<u8")]))
        cfg["device"] = device
        cfg["memory_budget"] = memory_budget
        h = _VP()
        _check(lib().rkfm_create(_p(cfg), C.byref(h)))
        self.h = h.value
        self.graphs = weakref.WeakSet()  # batch graphs, destroyed before the context
        _OPEN_CONTEXTS.add(self)

    def close(self):
        if getattr(self, "h", None):
            for g in list(getattr(self, "graphs", ())):
                g.destroy()
            lib().rkfm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        _check(lib().rkfm_sync(self.h))

    def timer_begin(self):
        _check(lib().rkfm_timer_begin(self.h))

    def timer_end(self):
        ms = _F()
        _check(lib().rkfm_timer_end(self.h, C.byref(ms)))
        return ms.value

    def copy(self, dst, src, nbytes):
        _check(lib().rkfm_device_copy(self.h, dst.ptr, src.ptr, nbytes))

    def gather16(self, dst, src, src_bytes, n):
        _check(lib().rkfm_device_gather16(self.h, dst.ptr, src.ptr, src_bytes, n))

    def array(self, shape, dtype):
        return DeviceArray(self, shape, dtype)

    def to_device(self, host):
        host = np.ascontiguousarray(host)
        d = DeviceArray(self, host.shape, host.dtype)
        d.upload(host)
        return d


class DeviceArray:
    """A device allocation made through the library (no framework needed)."""

    def __init__(self, ctx, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        h = _VP()
        _check(lib().rkfm_alloc(ctx.h, max(self.nbytes, 1), C.byref(h)))
        self.ptr = h.value

    def upload(self, host):
        host = np.ascontiguousarray(host, self.dtype)
        assert host.nbytes == self.nbytes
        _check(lib().rkfm_memcpy_h2d(self.ctx.h, self.ptr, host.ctypes.data, self.nbytes))

    def download(self):
        out = np.empty(self.shape, self.dtype)
        _check(lib().rkfm_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr, self.nbytes))
        return out

    def fill(self, byte):
        _check(lib().rkfm_memset(self.ctx.h, self.ptr, byte, self.nbytes))

    def free(self):
        if self.ptr:
            lib().rkfm_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.ctx.h:
                self.free()
        except Exception:
            pass


class GBuffer:
    """gbuffer.rs:14-58: position rgba32f, normal rgba16f, material r32u,
    motion rgba32f (+ optional debug and shadow planes), device-resident."""

    def __init__(self, ctx, width, height, debug=False, shadow=False):
        self.ctx, self.width, self.height = ctx, width, height
        n = width * height
        self.position = ctx.array((n, 4), np.float32)
        self.normal = ctx.array((n, 4), np.uint16)
        self.material = ctx.array((n,), np.uint32)
        self.motion = ctx.array((n, 4), np.float32)
        self.debug = ctx.array((n,), L.DEBUG_PX) if debug else None
        self.shadow = ctx.array((n,), np.float32) if shadow else None
        desc = np.zeros(1, L.GBUFFER_DESC)
        desc["position"] = self.position.ptr
        desc["normal"] = self.normal.ptr
        desc["material"] = self.material.ptr
        desc["motion"] = self.motion.ptr
        desc["debug"] = self.debug.ptr if debug else 0
        desc["shadow"] = self.shadow.ptr if shadow else 0
        desc["width"] = width
        desc["height"] = height
        self.desc = desc

    def read(self):
        out = dict(position=self.position.download(), normal=self.normal.download(),
                   material=self.material.download(), motion=self.motion.download())
        if self.debug is not None:
            out["debug"] = self.debug.download()
        if self.shadow is not None:
            out["shadow"] = self.shadow.download()
        return out


class GpuScene:
    """gpu_scene.rs:52-195: the group-0 inputs of the pass."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.num_objects = 0

    def upload_brick_pool(self, bricks, first_slot=0):
        b = np.ascontiguousarray(bricks, L.VOXEL_SAMPLE).reshape(-1)
        assert b.size % 512 == 0
        _check(lib().rkfm_upload_brick_pool(self.ctx.h, _p(b), b.size // 512, first_slot))

    def upload_brick_maps(self, entries, offset=0):
        e = np.ascontiguousarray(entries, np.uint32)
        _check(lib().rkfm_upload_brick_maps(self.ctx.h, _p(e), e.size, offset))

    def upload_objects(self, objects):
        o = np.ascontiguousarray(objects, L.GPU_OBJECT)
        _check(lib().rkfm_upload_objects(self.ctx.h, _p(o), o.size))
        self.num_objects = o.size

    def upload_bvh(self, nodes):
        n = np.ascontiguousarray(nodes, L.BVH_NODE) if nodes is not None else np.zeros(0, L.BVH_NODE)
        _check(lib().rkfm_upload_bvh(self.ctx.h, _p(n) if n.size else None, n.size))

    def update_camera(self, cam):
        c = np.ascontiguousarray(cam, L.CAMERA_UNIFORMS)
        _check(lib().rkfm_set_camera(self.ctx.h, _p(c)))

    def update_scene(self, su, ext=None):
        s = np.ascontiguousarray(su, L.SCENE_UNIFORMS)
        e = np.ascontiguousarray(ext, L.SCENE_EXT) if ext is not None else None
        _check(lib().rkfm_set_scene(self.ctx.h, _p(s), _p(e)))


class TileObjectCullPass:
    """tile_object_cull.rs:23-205 (16x16 tiles, <= 32 objects per tile)."""

    def __init__(self, ctx, width, height):
        self.ctx, self.width, self.height = ctx, width, height
        self.num_tiles_x = (width + 15) // 16
        self.num_tiles_y = (height + 15) // 16

    def dispatch(self):
        _check(lib().rkfm_build_tiles(self.ctx.h, self.width, self.height))

    def set_lists(self, idx, counts):
        i = np.ascontiguousarray(idx, np.uint32)
        c = np.ascontiguousarray(counts, np.uint32)
        _check(lib().rkfm_set_tiles(self.ctx.h, _p(i), _p(c), self.num_tiles_x, self.num_tiles_y))

    def clear(self):
        _check(lib().rkfm_set_tiles(self.ctx.h, None, None, 0, 0))

    def read(self):
        nt = self.num_tiles_x * self.num_tiles_y
        idx = np.zeros(nt * 32, np.uint32)
        counts = np.zeros(nt, np.uint32)
        _check(lib().rkfm_read_tiles(self.ctx.h, _p(idx), _p(counts), nt))
        return idx, counts


COARSE_VOXEL_SIZE = 0.32  # coarse_field.rs:20
MAX_COARSE_DIM = 128      # coarse_field.rs:23


class CoarseField:
    """coarse_field.rs:45-307: f16 unsigned distance to object AABBs."""

    def __init__(self, ctx, origin, extent, voxel_size=COARSE_VOXEL_SIZE):
        self.ctx = ctx
        self.origin = np.asarray(origin, np.float32)
        ext = np.asarray(extent, np.float32)
        self.dims = np.array([min(max(int(np.ceil(np.float32(e) / np.float32(voxel_size))), 1), MAX_COARSE_DIM)
                              for e in ext], np.uint32)
        self.voxel_size = np.float32(voxel_size)

    @classmethod
    def from_scene_aabbs(cls, ctx, aabbs, voxel_size=COARSE_VOXEL_SIZE, margin=1.0):
        """coarse_field.rs:211-237."""
        if len(aabbs) == 0:
            lo, hi = np.zeros(3, np.float32), np.full(3, voxel_size, np.float32)
        else:
            lo = np.min(np.array([a[0] for a in aabbs], np.float32), axis=0)
            hi = np.max(np.array([a[1] for a in aabbs], np.float32), axis=0)
        origin = lo - np.float32(margin)
        extent = (hi + np.float32(margin)) - origin
        return cls(ctx, origin, extent, voxel_size)

    def populate_and_upload(self):
        """populate (coarse_field.rs:240-266) on the GPU + upload uniforms
        relative to the scene camera (coarse_field.rs:279-307)."""
        _check(lib().rkfm_build_coarse_field(self.ctx.h, _p(self.origin), _p(self.dims), self.voxel_size))

    def read(self):
        n = int(np.prod(self.dims))
        cells = np.zeros(n, np.uint16)
        u = np.zeros(1, L.COARSE_FIELD_UNIFORMS)
        _check(lib().rkfm_read_coarse_field(self.ctx.h, _p(cells), n, _p(u)))
        return cells, u[0]

    def upload_cells(self, cells, uniforms):
        c = np.ascontiguousarray(cells, np.uint16)
        u = np.ascontiguousarray(uniforms, L.COARSE_FIELD_UNIFORMS)
        _check(lib().rkfm_set_coarse_field(self.ctx.h, _p(c), _p(u)))

    def clear(self):
        _check(lib().rkfm_set_coarse_field(self.ctx.h, None, None))


class RayMarchPass:
    """ray_march.rs:29-118."""

    def __init__(self, ctx, scene=None, gbuffer=None, tile_cull=None, coarse_field=None):
        self.ctx = ctx

    def recreate_pipeline(self, *_args):
        """Hot reload has no meaning for ahead-of-time compiled kernels: no-op."""

    def dispatch(self, scene, gbuffer, tile_cull=None, coarse_field=None, stream=None):
        _check(lib().rkfm_dispatch(self.ctx.h, gbuffer.desc.ctypes.data, stream))


def voxelize_sphere(radius, margin, voxel_size, object_index=0, mat_seed=0, material=0, first_slot=0):
    """Host voxelize_sdf of |p| - r over +-(r + margin). Returns (dims, map, bricks)."""
    dims = np.zeros(3, np.uint32)
    nb = np.zeros(1, np.uint64)
    _check(lib().rkfm_voxelize_sphere(radius, margin, voxel_size, object_index, mat_seed, material, first_slot,
                                      _p(dims), None, None, _p(nb)))
    m = np.empty(int(np.prod(dims)), np.uint32)
    b = np.empty(int(nb[0]) * 512, L.VOXEL_SAMPLE)
    _check(lib().rkfm_voxelize_sphere(radius, margin, voxel_size, object_index, mat_seed, material, first_slot,
                                      _p(dims), _p(m), _p(b), _p(nb)))
    return dims, m, b


def voxelize_prim(prim, margin, voxel_size, flags=0, object_index=0, mat_seed=0, material=0, first_slot=0):
    """Host voxelisation of one primitive record (rkfm_voxelize_prim; flags
    VOX_GEOMETRY_FIRST / VOX_BLEND). Returns (dims, map, bricks)."""
    pr = np.ascontiguousarray(prim, L.PRIM)
    dims = np.zeros(3, np.uint32)
    nb = np.zeros(1, np.uint64)
    args = (margin, voxel_size, flags, object_index, mat_seed, material, first_slot)
    _check(lib().rkfm_voxelize_prim(_p(pr), *args, _p(dims), None, None, _p(nb)))
    m = np.empty(int(np.prod(dims)), np.uint32)
    b = np.empty(int(nb[0]) * 512, L.VOXEL_SAMPLE)
    _check(lib().rkfm_voxelize_prim(_p(pr), *args, _p(dims), _p(m), _p(b), _p(nb)))
    return dims, m, b


def voxelize_test_cross():
    dims = np.zeros(3, np.uint32)
    nb = np.zeros(1, np.uint64)
    _check(lib().rkfm_voxelize_test_cross(_p(dims), None, None, _p(nb)))
    m = np.empty(int(np.prod(dims)), np.uint32)
    b = np.empty(int(nb[0]) * 512, L.VOXEL_SAMPLE)
    _check(lib().rkfm_voxelize_test_cross(_p(dims), _p(m), _p(b), _p(nb)))
    return dims, m, b


def camera_look_at(position, target, fov_degrees, width, height):
    out = np.zeros(1, L.CAMERA_UNIFORMS)
    pos = np.asarray(position, np.float32)
    tgt = np.asarray(target, np.float32)
    _check(lib().rkfm_camera_look_at(_p(pos), _p(tgt), fov_degrees, width, height, _p(out)))
    return out[0]


def build_bvh(objects):
    o = np.ascontiguousarray(objects, L.GPU_OBJECT)
    nodes = np.zeros(max(2 * o.size - 1, 1), L.BVH_NODE)
    n = np.zeros(1, np.uint32)
    _check(lib().rkfm_build_bvh(_p(o), o.size, _p(nodes), _p(n)))
    return nodes[: int(n[0])]


class SparseScene:
    """A resident two-level sparse scene (see include/rkf_march.h)."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.h = handle
        info = np.zeros(1, L.SPARSE_INFO)
        _check(lib().rkfm_sparse_info_get(self.h, _p(info)))
        self.info = info[0]

    @classmethod
    def build(cls, ctx, prims, voxel_size, bounds_min, bounds_max, smin_k=0.0, mat_seed=0, max_bricks=0):
        p = np.ascontiguousarray(prims, L.PRIM)
        prm = np.zeros(1, L.SPARSE_PARAMS)
        prm["voxel_size"] = voxel_size
        prm["smin_k"] = smin_k
        prm["mat_seed"] = mat_seed
        prm["bounds_min"] = bounds_min
        prm["bounds_max"] = bounds_max
        prm["max_bricks"] = max_bricks
        h = _VP()
        _check(lib().rkfm_sparse_build(ctx.h, _p(p), p.size, _p(prm), C.byref(h)))
        return cls(ctx, h.value)

    @classmethod
    def from_dense(cls, ctx, dims, brick_map, bricks, voxel_size, origin):
        d = np.asarray(dims, np.uint32)
        m = np.ascontiguousarray(brick_map, np.uint32)
        b = np.ascontiguousarray(bricks, L.VOXEL_SAMPLE).reshape(-1)
        o = np.asarray(origin, np.float32)
        h = _VP()
        _check(lib().rkfm_sparse_from_dense(ctx.h, _p(m), _p(d), _p(b), b.size // 512, voxel_size, _p(o),
                                            C.byref(h)))
        return cls(ctx, h.value)

    @classmethod
    def from_objects(cls, ctx, voxel_size, bounds_min, bounds_max):
        """Resident scene resampled from the reference-layout scene uploaded
        to ctx (rkfm_sparse_from_objects)."""
        lo = np.asarray(bounds_min, np.float32)
        hi = np.asarray(bounds_max, np.float32)
        h = _VP()
        _check(lib().rkfm_sparse_from_objects(ctx.h, voxel_size, _p(lo), _p(hi), C.byref(h)))
        return cls(ctx, h.value)

    def download_aux(self):
        """The aux plane (n_bricks * 512 u16) or None when the scene has none."""
        present = np.zeros(1, np.uint32)
        _check(lib().rkfm_sparse_download_aux(self.ctx.h, self.h, None, _p(present)))
        if not present[0]:
            return None
        aux = np.empty(int(self.info["n_bricks"]) * 512, np.uint16)
        _check(lib().rkfm_sparse_download_aux(self.ctx.h, self.h, _p(aux), _p(present)))
        return aux

    @classmethod
    def from_rkf(cls, ctx, data, lod=0, origin=None):
        """Resident scene from a .rkf v2 asset (bytes); origin None = centred grid."""
        buf = np.frombuffer(bytes(data), np.uint8)
        o = None if origin is None else np.asarray(origin, np.float32)
        h = _VP()
        _check(lib().rkfm_sparse_from_rkf(ctx.h, buf.ctypes.data, buf.size, lod, _p(o), C.byref(h)))
        return cls(ctx, h.value)

    def destroy(self):
        if self.h:
            _check(lib().rkfm_sparse_destroy(self.ctx.h, self.h))
            self.h = None

    def download(self):
        i = self.info
        nc = int(np.prod(i["cells"]))
        l2 = np.zeros(nc, np.uint32)
        l1 = np.zeros(int(i["n_blocks"]) * 512, np.uint32)
        dist = np.zeros(int(i["n_bricks"]) * 512, np.uint16)
        mat = np.zeros(int(i["n_bricks"]) * 512, np.uint8)
        _check(lib().rkfm_sparse_download(self.ctx.h, self.h, _p(l2), _p(l1), _p(dist), _p(mat)))
        return l2, l1, dist, mat

    def download_free(self):
        """(free_idx, free_bits, oct, brick_ext) as in rkfm_sparse_download_free."""
        nc = int(np.prod(self.info["cells"]))
        rows = np.zeros(1, np.uint64)
        _check(lib().rkfm_sparse_download_free(self.ctx.h, self.h, None, None, None, None, _p(rows)))
        fidx = np.zeros(nc, np.uint32)
        bits = np.zeros(max(int(rows[0]) * 8, 1), np.uint64)
        oct_ = np.zeros(nc, np.uint32)
        bext = np.zeros(max(int(rows[0]) * 2048, 1), np.uint8)
        _check(lib().rkfm_sparse_download_free(self.ctx.h, self.h, _p(fidx), _p(bits), _p(oct_), _p(bext),
                                               _p(rows)))
        return fidx, bits, oct_, bext

    def march(self, cam, su, ext, gbuffer, variant=0, stream=None):
        c = np.ascontiguousarray(cam, L.CAMERA_UNIFORMS)
        s = np.ascontiguousarray(su, L.SCENE_UNIFORMS)
        e = np.ascontiguousarray(ext, L.SCENE_EXT) if ext is not None else None
        _check(lib().rkfm_sparse_march(self.ctx.h, self.h, _p(c), _p(s), _p(e), gbuffer.desc.ctypes.data,
                                       variant, stream))

    def march_noalloc(self, cam, su, ext, gbuffer, variant=0, stream=None):
        """rkfm_sparse_march_noalloc: the same march with its scratch in the
        library's fixed per-device array (no device allocation)."""
        c = np.ascontiguousarray(cam, L.CAMERA_UNIFORMS)
        s = np.ascontiguousarray(su, L.SCENE_UNIFORMS)
        e = np.ascontiguousarray(ext, L.SCENE_EXT) if ext is not None else None
        _check(lib().rkfm_sparse_march_noalloc(self.ctx.h, self.h, _p(c), _p(s), _p(e), gbuffer.desc.ctypes.data,
                                               variant, stream))

    def tile_intervals(self, cam, su, width, height):
        """(t_lo, t_hi) per 8x8 tile (rkfm_sparse_tile_intervals), shape
        (tiles_y, tiles_x, 2) float32."""
        c = np.ascontiguousarray(cam, L.CAMERA_UNIFORMS)
        s = np.ascontiguousarray(su, L.SCENE_UNIFORMS)
        tx, ty = (width + 7) // 8, (height + 7) // 8
        d = self.ctx.array((ty, tx, 2), np.float32)
        _check(lib().rkfm_sparse_tile_intervals(self.ctx.h, self.h, _p(c), _p(s), width, height, d.ptr, None))
        self.ctx.sync()
        return d.download()

    def tile_order(self, cam, su, width, height):
        """(counts, order) of a culled view (rkfm_sparse_tile_order): cells
        binned into each 8x8 tile, shape (tiles_y, tiles_x), and the tiles'
        launch order (flat tile indices)."""
        c = np.ascontiguousarray(cam, L.CAMERA_UNIFORMS)
        s = np.ascontiguousarray(su, L.SCENE_UNIFORMS)
        tx, ty = (width + 7) // 8, (height + 7) // 8
        cd = self.ctx.array((ty, tx), np.uint32)
        od = self.ctx.array((tx * ty,), np.uint32)
        _check(lib().rkfm_sparse_tile_order(self.ctx.h, self.h, _p(c), _p(s), width, height, cd.ptr, od.ptr, None))
        self.ctx.sync()
        return cd.download(), od.download()

    def edit(self, ops, report=True, stream=None):
        """Applies CSG edits in order (rkfm_sparse_edit); with report=True
        waits and returns the sorted slots of the modified bricks."""
        o = np.ascontiguousarray(np.atleast_1d(ops), L.EDIT)
        if not report:
            _check(lib().rkfm_sparse_edit(self.ctx.h, self.h, o.ctypes.data, o.size, None, 0, None, stream))
            return None
        cap = max(int(self.info["n_bricks"]), 1)
        out = np.zeros(cap, np.uint32)
        n = np.zeros(1, np.uint64)
        _check(lib().rkfm_sparse_edit(self.ctx.h, self.h, o.ctypes.data, o.size, out.ctypes.data, cap,
                                      n.ctypes.data, stream))
        return out[: int(n[0])]

    def contacts(self, bodies_dev, contacts_dev, counts_dev, n, stream=None):
        """generate_sdf_contacts for n bodies (device arrays: BODY, CONTACT x 16
        per body, u32 counts)."""
        _check(lib().rkfm_sparse_contacts(self.ctx.h, self.h, _p(bodies_dev), n, _p(contacts_dev), _p(counts_dev),
                                          stream))

    def query(self, xyz_dev, out_dev, n, stream=None):
        _check(lib().rkfm_sparse_query_points(self.ctx.h, self.h, _p(xyz_dev), n, _p(out_dev), stream))

    def sphere_cast(self, casts_dev, out_dev, n, stream=None):
        """rkfm_sparse_sphere_cast: casts_dev SPHERE_CAST records -> out_dev CAST_RESULT records."""
        _check(lib().rkfm_sparse_sphere_cast(self.ctx.h, self.h, _p(casts_dev), n, _p(out_dev), stream))


class Batch:
    """A set of uneven batches (views, shadow rays, point queries) over
    resident sparse scenes, run in one launch (rkfm_batch_run) or captured
    into a HIP graph. Keeps its device buffers alive."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.items = []
        self.keep = []

    def add_view(self, scene, cam, su, ext, gbuffer):
        it = np.zeros((), L.BATCH_ITEM)
        it["kind"] = L.BATCH_VIEW
        it["scene"] = scene.h
        it["camera"] = cam
        it["su"] = su
        it["ext"] = ext
        it["out"] = gbuffer.desc[0]
        self.items.append(it)
        self.keep.append(gbuffer)
        return len(self.items) - 1

    def add_shadow(self, scene, origins_dev, vis_dev, n, ext):
        it = np.zeros((), L.BATCH_ITEM)
        it["kind"] = L.BATCH_SHADOW
        it["n"] = n
        it["scene"] = scene.h
        it["ext"] = ext
        it["origins"] = origins_dev.ptr
        it["visibility"] = vis_dev.ptr
        self.items.append(it)
        self.keep += [origins_dev, vis_dev]
        return len(self.items) - 1

    def add_points(self, scene, xyz_dev, out_dev, n):
        it = np.zeros((), L.BATCH_ITEM)
        it["kind"] = L.BATCH_POINTS
        it["n"] = n
        it["scene"] = scene.h
        it["xyz"] = xyz_dev.ptr
        it["results"] = out_dev.ptr
        self.items.append(it)
        self.keep += [xyz_dev, out_dev]
        return len(self.items) - 1

    def add_spheres(self, scene, casts_dev, out_dev, n):
        it = np.zeros((), L.BATCH_ITEM)
        it["kind"] = L.BATCH_SPHERES
        it["n"] = n
        it["scene"] = scene.h
        it["casts"] = casts_dev.ptr
        it["cast_results"] = out_dev.ptr
        self.items.append(it)
        self.keep += [casts_dev, out_dev]
        return len(self.items) - 1

    def array(self):
        return np.ascontiguousarray(np.array(self.items, L.BATCH_ITEM))

    def work_items(self):
        a = self.array()
        return int(lib().rkfm_batch_work_items(a.ctypes.data, a.size))

    def run(self, tuning=None, stream=None):
        a = self.array()
        t = None
        if tuning is not None:
            t = np.zeros(1, L.BATCH_TUNING)
            t[0] = tuning
        _check(lib().rkfm_batch_run(self.ctx.h, a.ctypes.data, a.size, _p(t), stream))

    def autotune(self, reps=3):
        a = self.array()
        t = np.zeros(1, L.BATCH_TUNING)
        ms = _F()
        _check(lib().rkfm_batch_autotune(self.ctx.h, a.ctypes.data, a.size, reps, t.ctypes.data, C.byref(ms)))
        return t[0], ms.value

    def graph(self, tuning=None):
        a = self.array()
        t = None
        if tuning is not None:
            t = np.zeros(1, L.BATCH_TUNING)
            t[0] = tuning
        h = _VP()
        _check(lib().rkfm_batch_graph_create(self.ctx.h, a.ctypes.data, a.size, _p(t), C.byref(h)))
        return BatchGraph(self, h.value)


class BatchGraph:
    def __init__(self, batch, handle):
        self.batch = batch
        self.h = handle
        batch.ctx.graphs.add(self)  # a graph refers to its context: destroyed first

    def launch(self, stream=None):
        _check(lib().rkfm_batch_graph_launch(self.h, stream))

    def destroy(self):
        if self.h:
            lib().rkfm_batch_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class PinnedArray:
    """Pinned host memory (rkfm_host_alloc) viewed as a numpy array
    (``.a``); copies between it and the device run asynchronously."""

    def __init__(self, shape, dtype):
        self.dtype = np.dtype(dtype)
        self.shape = tuple(np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        count = int(np.prod(self.shape))
        nbytes = max(count * self.dtype.itemsize, 1)
        h = _VP()
        _check(lib().rkfm_host_alloc(nbytes, C.byref(h)))
        self.ptr = h.value
        buf = (C.c_ubyte * nbytes).from_address(self.ptr)
        self.a = np.frombuffer(buf, dtype=self.dtype, count=count).reshape(self.shape)

    def free(self):
        if self.ptr:
            self.a = None
            lib().rkfm_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Server:
    """rkfm_server: resident scenes, memory-budget admission, continuous
    batching (see include/rkf_march.h). Request builders return REQUEST
    records; submit() keeps the host arrays a request points at alive until
    release()."""

    def __init__(self, ctx, scenes, memory_budget=1 << 30, max_requests=0, rounds_in_flight=0, max_round_work=0,
                 max_round_requests=0, policy=L.SERVE_BATCHED, tuning=None, max_pending=0):
        self.ctx = ctx
        self.scenes = list(scenes)
        handles = np.array([s.h for s in self.scenes], np.uint64)
        cfg = np.zeros(1, L.SERVER_CONFIG)
        cfg["memory_budget"] = memory_budget
        cfg["max_requests"] = max_requests
        cfg["rounds_in_flight"] = rounds_in_flight
        cfg["max_round_work"] = max_round_work
        cfg["max_round_requests"] = max_round_requests
        cfg["policy"] = policy
        cfg["max_pending"] = max_pending
        if tuning is not None:
            cfg["tuning"] = tuning
        h = _VP()
        _check(lib().rkfm_server_create(ctx.h, handles.ctypes.data, handles.size, cfg.ctypes.data, C.byref(h)))
        self.h = h.value
        self._keep = {}
        _OPEN_SERVERS.add(self)

    # request builders -------------------------------------------------------
    @staticmethod
    def view_request(scene, cam, su, ext, width, height, keep_on_device=True, host=None, tag=0):
        """host: dict of PinnedArray planes (position/normal/material/motion/shadow)."""
        r = np.zeros((), L.REQUEST)
        r["kind"] = L.BATCH_VIEW
        r["scene"] = scene
        r["flags"] = L.REQ_KEEP_ON_DEVICE if keep_on_device else 0
        r["width"], r["height"] = width, height
        r["camera"], r["su"], r["ext"] = cam, su, ext
        r["tag"] = tag
        for k, arr in (host or {}).items():
            r["view_out"][k] = arr.ptr
        return r

    @staticmethod
    def points_request(scene, xyz, out, n=None, tag=0):
        r = np.zeros((), L.REQUEST)
        r["kind"] = L.BATCH_POINTS
        r["scene"] = scene
        r["n"] = xyz.shape[0] if n is None else n
        r["in"], r["out"] = xyz.ptr, out.ptr
        r["tag"] = tag
        return r

    @staticmethod
    def cast_request(scene, casts, out, n=None, tag=0):
        """casts: PinnedArray of SPHERE_CAST records, out: PinnedArray of CAST_RESULT."""
        r = np.zeros((), L.REQUEST)
        r["kind"] = L.BATCH_SPHERES
        r["scene"] = scene
        r["n"] = casts.shape[0] if n is None else n
        r["in"], r["out"] = casts.ptr, out.ptr
        r["tag"] = tag
        return r

    @staticmethod
    def shadow_request(scene, origins, vis, ext, n=None, tag=0):
        r = np.zeros((), L.REQUEST)
        r["kind"] = L.BATCH_SHADOW
        r["scene"] = scene
        r["n"] = origins.shape[0] if n is None else n
        r["ext"] = ext
        r["in"], r["out"] = origins.ptr, vis.ptr
        r["tag"] = tag
        return r

    # calls ------------------------------------------------------------------
    def submit(self, req, keep=()):
        r = np.ascontiguousarray(req, L.REQUEST).reshape(1)
        i = _U64()
        _check(lib().rkfm_server_submit(self.h, r.ctypes.data, C.byref(i)))
        self._keep[i.value] = keep
        return i.value

    def status(self, rid):
        st = np.zeros(1, L.REQUEST_STATUS)
        _check(lib().rkfm_server_status(self.h, rid, st.ctypes.data))
        return st[0]

    def wait(self, rid, timeout=-1.0):
        st = np.zeros(1, L.REQUEST_STATUS)
        _check(lib().rkfm_server_wait(self.h, rid, timeout, st.ctypes.data))
        return st[0]

    def view_planes(self, rid):
        g = np.zeros(1, L.GBUFFER_DESC)
        _check(lib().rkfm_server_view_planes(self.h, rid, g.ctypes.data))
        return g[0]

    def read_view(self, rid):
        """Host copies of a finished view kept on the device."""
        g = self.view_planes(rid)
        n = int(g["width"]) * int(g["height"])
        out = {}
        for k, shape, dt in (("position", (n, 4), np.float32), ("normal", (n, 4), np.uint16),
                             ("material", (n,), np.uint32), ("motion", (n, 4), np.float32),
                             ("shadow", (n,), np.float32)):
            if int(g[k]) == 0:
                continue
            a = np.empty(shape, dt)
            _check(lib().rkfm_memcpy_d2h(self.ctx.h, a.ctypes.data, int(g[k]), a.nbytes))
            out[k] = a
        return out

    def release(self, rid):
        _check(lib().rkfm_server_release(self.h, rid))
        self._keep.pop(rid, None)

    def stats(self):
        st = np.zeros(1, L.SERVER_STATS)
        _check(lib().rkfm_server_stats_get(self.h, st.ctypes.data))
        return st[0]

    def now(self):
        return lib().rkfm_server_now(self.h)

    def replay(self, reqs, arrivals):
        """Native replay of an arrival trace; returns (latency seconds per
        request, NaN if failed; wall seconds to the last completion)."""
        r = np.ascontiguousarray(reqs, L.REQUEST)
        a = np.ascontiguousarray(arrivals, np.float64)
        lat = np.full(r.size, np.nan)
        wall = C.c_double()
        _check(lib().rkfm_server_replay(self.h, r.ctypes.data, a.ctypes.data, r.size, lat.ctypes.data,
                                        C.byref(wall)))
        return lat, wall.value

    def close(self):
        if getattr(self, "h", None):
            lib().rkfm_server_destroy(self.h)
            self.h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rkf_read_header(data):
    """(.rkf header record, LOD table) of an asset."""
    buf = np.frombuffer(bytes(data), np.uint8)
    h = np.zeros(1, L.RKF_HEADER)
    _check(lib().rkfm_rkf_read_header(buf.ctypes.data, buf.size, h.ctypes.data, None, 0))
    lods = np.zeros(max(int(h[0]["lod_count"]), 1), L.RKF_LOD)
    _check(lib().rkfm_rkf_read_header(buf.ctypes.data, buf.size, h.ctypes.data, lods.ctypes.data, lods.size))
    return h[0], lods[: int(h[0]["lod_count"])]


def rkf_decode_lod(data, lod=0):
    """(dims, brick map, bricks) of one LOD of an asset."""
    buf = np.frombuffer(bytes(data), np.uint8)
    h, lods = rkf_read_header(data)
    if lod >= len(lods):
        m = np.zeros(1, np.uint32)
        _check(lib().rkfm_rkf_decode_lod(buf.ctypes.data, buf.size, lod, m.ctypes.data, None))
    e = lods[lod]
    dims = np.array(e["brick_dims"], np.uint32)
    # validate (decode into nothing) before sizing arrays from the LOD table,
    # whose fields a damaged file can set to anything
    _check(lib().rkfm_rkf_decode_lod(buf.ctypes.data, buf.size, lod, None, None))
    nmap = int(dims[0]) * int(dims[1]) * int(dims[2])
    nb = int(e["brick_count"]) * 512
    m = np.zeros(max(nmap, 1), np.uint32)
    b = np.zeros(max(nb, 1), L.VOXEL_SAMPLE)
    _check(lib().rkfm_rkf_decode_lod(buf.ctypes.data, buf.size, lod, m.ctypes.data, b.ctypes.data))
    return dims, m[:nmap], b[:nb]


def rkf_encode(levels, aabb_min=(0, 0, 0), aabb_max=(1, 1, 1), material_ids=(), analytical=(0, (0, 0, 0, 0))):
    """save_object: levels = [(voxel_size, dims, map, bricks)], map holding
    slots (renumbered in order of first appearance), bricks in that order."""
    hdr = np.zeros(1, L.RKF_HEADER)
    hdr["aabb_min"], hdr["aabb_max"] = aabb_min, aabb_max
    mats = list(material_ids)[:16]
    hdr["material_count"] = len(mats)
    hdr[0]["material_ids"][: len(mats)] = mats
    hdr["analytical_type"] = analytical[0]
    hdr["analytical_params"] = analytical[1]
    src = np.zeros(len(levels), L.RKF_LOD_SRC)
    keep = []
    for i, (vs, dims, m, b) in enumerate(levels):
        m = np.ascontiguousarray(m, np.uint32)
        b = np.ascontiguousarray(b, L.VOXEL_SAMPLE).reshape(-1)
        keep += [m, b]
        src[i]["voxel_size"] = vs
        src[i]["dims"] = dims
        src[i]["map"] = m.ctypes.data if m.size else 0
        src[i]["bricks"] = b.ctypes.data if b.size else 0
        src[i]["brick_count"] = b.size // 512
    n = np.zeros(1, np.uint64)
    _check(lib().rkfm_rkf_encode(hdr.ctypes.data, src.ctypes.data, len(levels), None, 0, n.ctypes.data))
    out = np.zeros(int(n[0]), np.uint8)
    _check(lib().rkfm_rkf_encode(hdr.ctypes.data, src.ctypes.data, len(levels), out.ctypes.data, out.size,
                                 n.ctypes.data))
    return out.tobytes()

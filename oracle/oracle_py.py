"""ctypes front-end of the CPU oracle (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker. The product package never
imports this module.
"""
import ctypes as C
import os

import numpy as np

from rkifield_amd import layouts as L

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

RAY_RESULT = np.dtype([
    ("hit", "<i4"), ("flags", "<u4"),
    ("t", "<f8"),
    ("pos", "<f8", (3,)),
    ("normal", "<f8", (3,)),
    ("motion", "<f8", (2,)),
    ("sdf_at_hit", "<f8"),
    ("packed_material", "<u4"),
    ("material_id", "<u4"),
    ("obj_idx", "<u4"),
    ("object_id", "<u4"),
    ("brick_slot", "<u4"),
    ("voxel_index", "<u4"),
    ("cell_index", "<u4"),
    ("steps", "<u4"),
    ("evals", "<u4"),
    ("shadow_steps", "<u4"),
    ("shadow_evals", "<u4"), ("path", "<u4"),
    ("shadow", "<f8"),
])
assert RAY_RESULT.itemsize == 144


class RefScene(C.Structure):
    _fields_ = [
        ("pool", C.c_void_p), ("n_pool_voxels", C.c_uint64),
        ("maps", C.c_void_p), ("n_maps", C.c_uint64),
        ("objects", C.c_void_p),
        ("bvh", C.c_void_p), ("n_bvh", C.c_uint32),
        ("tile_idx", C.c_void_p), ("tile_counts", C.c_void_p),
        ("tiles_x", C.c_uint32), ("tiles_y", C.c_uint32),
        ("coarse", C.c_void_p),
        ("coarse_u", C.c_uint8 * 48),
        ("cam", C.c_uint8 * 144),
        ("su", C.c_uint8 * 16),
        ("ext", C.c_uint8 * 48),
        ("width", C.c_uint32), ("height", C.c_uint32),
    ]


class SparseScene(C.Structure):
    _fields_ = [
        ("origin", C.c_float * 3), ("voxel_size", C.c_float),
        ("cells", C.c_uint32 * 3),
        ("l2", C.c_void_p), ("l1", C.c_void_p), ("dist", C.c_void_p), ("mat", C.c_void_p),
        ("free_l2", C.c_void_p), ("oct", C.c_void_p), ("bext", C.c_void_p),
        ("aux", C.c_void_p),
    ]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liborc.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liborc.so is not built (run __graft_entry__.build())")
        _LIB = C.CDLL(path)
        _LIB.orc_f32_to_f16.restype = C.c_uint16
        _LIB.orc_f32_to_f16.argtypes = [C.c_float]
        _LIB.orc_f16_to_f32.restype = C.c_float
        _LIB.orc_f16_to_f32.argtypes = [C.c_uint16]
        _LIB.orc_brick_index.restype = C.c_int64
        _LIB.orc_map_flatten.restype = C.c_int64
        _LIB.orc_pool_allocate.restype = C.c_int64
        _LIB.orc_sample_brick_trilinear.restype = C.c_float
        _LIB.orc_sample_brick_nearest_material.restype = C.c_uint32
        _LIB.orc_sparse_sample.restype = C.c_double
        _LIB.orc_sparse_sample.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        _LIB.orc_prim_eval.restype = C.c_float
        _LIB.orc_smin.restype = C.c_float
        _LIB.orc_smin.argtypes = [C.c_float, C.c_float, C.c_float]
        _LIB.orc_voxelize_sdf.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_uint32,
                                          C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        _LIB.orc_voxelize_geometry.argtypes = _LIB.orc_voxelize_sdf.argtypes
    return _LIB


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None else C.c_void_p(0)


def f32_to_f16_bits(x):
    x = np.ascontiguousarray(x, np.float32).ravel()
    out = np.empty(x.size, np.uint16)
    lib().orc_f32_to_f16_array(_p(x), _p(out), C.c_uint64(x.size))
    return out


def f16_bits_to_f32(h):
    h = np.ascontiguousarray(h, np.uint16).ravel()
    out = np.empty(h.size, np.float32)
    lib().orc_f16_to_f32_array(_p(h), _p(out), C.c_uint64(h.size))
    return out


def voxel_new(d, mat, blend=0):
    lib().orc_voxel_new.restype = C.c_uint64
    lib().orc_voxel_new.argtypes = [C.c_float, C.c_uint16, C.c_uint8]
    v = lib().orc_voxel_new(d, mat, blend)
    return int(v) & 0xFFFFFFFF, int(v) >> 32


def brick_index(x, y, z):
    return lib().orc_brick_index(C.c_uint32(x), C.c_uint32(y), C.c_uint32(z))


def map_flatten(dims, bx, by, bz):
    d = np.asarray(dims, np.uint32)
    return lib().orc_map_flatten(_p(d), C.c_uint32(bx), C.c_uint32(by), C.c_uint32(bz))


def sample_brick_trilinear(brick, pos):
    b = np.ascontiguousarray(brick, L.VOXEL_SAMPLE)
    p = np.asarray(pos, np.float32)
    return lib().orc_sample_brick_trilinear(_p(b), _p(p))


def sample_brick_nearest_material(brick, pos):
    b = np.ascontiguousarray(brick, L.VOXEL_SAMPLE)
    p = np.asarray(pos, np.float32)
    return lib().orc_sample_brick_nearest_material(_p(b), _p(p))


def pool_allocate_sequence(capacity, count):
    class Pool(C.Structure):
        _fields_ = [("capacity", C.c_uint32), ("free_list", C.c_void_p), ("n_free", C.c_uint32)]
    p = Pool()
    lib().orc_pool_init(C.byref(p), C.c_uint32(capacity))
    out = [lib().orc_pool_allocate(C.byref(p)) for _ in range(count)]
    lib().orc_pool_free(C.byref(p))
    return out


def voxelize_test_cross(first_slot=0):
    mn = np.zeros(3, np.float32)
    mx = np.zeros(3, np.float32)
    lib().orc_test_cross_aabb(_p(mn), _p(mx))
    dims = np.zeros(3, np.uint32)
    nb = C.c_uint64(0)
    fn = C.cast(lib().orc_test_cross_sdf, C.c_void_p).value
    lib().orc_voxelize_sdf(fn, None, _p(mn), _p(mx), C.c_float(0.03), C.c_uint32(first_slot),
                           _p(dims), None, None, C.byref(nb))
    m = np.empty(int(np.prod(dims)), np.uint32)
    b = np.empty(nb.value * 512, L.VOXEL_SAMPLE)
    lib().orc_voxelize_sdf(fn, None, _p(mn), _p(mx), C.c_float(0.03), C.c_uint32(first_slot),
                           _p(dims), _p(m), _p(b), C.byref(nb))
    return dims, m, b


def rkf_load(data: bytes):
    """Parse a .rkf v2 file (LOD 0): returns (header dict, map, bricks)."""
    class Hdr(C.Structure):
        _fields_ = [("version", C.c_uint32), ("lod_count", C.c_uint32), ("material_count", C.c_uint32),
                    ("aabb_min", C.c_float * 3), ("aabb_max", C.c_float * 3),
                    ("material_ids", C.c_uint16 * 16), ("voxel_size", C.c_float),
                    ("brick_count", C.c_uint32), ("dims", C.c_uint32 * 3),
                    ("data_offset", C.c_uint64), ("compressed_size", C.c_uint32),
                    ("uncompressed_size", C.c_uint32)]
    buf = np.frombuffer(data, np.uint8)
    h = Hdr()
    rc = lib().orc_rkf_load(_p(buf), C.c_uint64(buf.size), C.byref(h), None, None)
    if rc != 0:
        raise ValueError(f"rkf header parse failed: {rc}")
    dims = np.array(h.dims[:], np.uint32)
    m = np.empty(int(np.prod(dims)), np.uint32)
    b = np.empty(h.brick_count * 512, L.VOXEL_SAMPLE)
    rc = lib().orc_rkf_load(_p(buf), C.c_uint64(buf.size), C.byref(h), _p(m), _p(b))
    if rc != 0:
        raise ValueError(f"rkf decode failed: {rc}")
    hdr = dict(version=h.version, lod_count=h.lod_count, material_count=h.material_count,
               aabb_min=np.array(h.aabb_min[:], np.float32), aabb_max=np.array(h.aabb_max[:], np.float32),
               material_ids=list(h.material_ids[:h.material_count]), voxel_size=h.voxel_size,
               brick_count=h.brick_count, dims=dims, compressed_size=h.compressed_size,
               uncompressed_size=h.uncompressed_size)
    return hdr, m, b


def coarse_populate(origin, dims, voxel_size, objects):
    o = np.asarray(origin, np.float32)
    d = np.asarray(dims, np.uint32)
    objs = np.ascontiguousarray(objects, L.GPU_OBJECT)
    cells = np.empty(int(np.prod(d)), np.uint16)
    lib().orc_coarse_populate(_p(o), _p(d), C.c_float(voxel_size), _p(objs), C.c_uint32(objs.size), _p(cells))
    return cells


def tile_cull(cam, su, objects, width, height):
    ntx, nty = (width + 15) // 16, (height + 15) // 16
    idx = np.zeros(ntx * nty * 32, np.uint32)
    counts = np.zeros(ntx * nty, np.uint32)
    cam = np.ascontiguousarray(cam, L.CAMERA_UNIFORMS)
    su = np.ascontiguousarray(su, L.SCENE_UNIFORMS)
    objs = np.ascontiguousarray(objects, L.GPU_OBJECT)
    lib().orc_tile_cull(_p(cam), _p(su), _p(objs), C.c_uint32(width), C.c_uint32(height), _p(idx), _p(counts))
    return idx, counts


def _bytes(arr, n):
    a = np.ascontiguousarray(arr)
    return (C.c_uint8 * n).from_buffer_copy(a.tobytes())


class RefInputs:
    """Holds numpy arrays alive and the C struct pointing at them."""

    def __init__(self, pool, maps, objects, cam, su, ext, width, height, bvh=None,
                 tile_idx=None, tile_counts=None, coarse=None, coarse_u=None):
        self.keep = [np.ascontiguousarray(x) if x is not None else None
                     for x in (pool, maps, objects, bvh, tile_idx, tile_counts, coarse)]
        pool, maps, objects, bvh, tile_idx, tile_counts, coarse = self.keep
        s = RefScene()
        s.pool = _p(pool).value
        s.n_pool_voxels = pool.size
        s.maps = _p(maps).value
        s.n_maps = maps.size
        s.objects = _p(objects).value
        s.bvh = _p(bvh).value if bvh is not None else None
        s.n_bvh = bvh.size if bvh is not None else 0
        s.tile_idx = _p(tile_idx).value if tile_idx is not None else None
        s.tile_counts = _p(tile_counts).value if tile_counts is not None else None
        ntx, nty = (width + 15) // 16, (height + 15) // 16
        s.tiles_x = ntx if tile_counts is not None else 0
        s.tiles_y = nty if tile_counts is not None else 0
        s.coarse = _p(coarse).value if coarse is not None else None
        if coarse_u is not None:
            s.coarse_u = _bytes(coarse_u, 48)
        s.cam = _bytes(cam, 144)
        s.su = _bytes(su, 16)
        s.ext = _bytes(ext, 48)
        s.width = width
        s.height = height
        self.s = s
        self.width, self.height = width, height


def ref_trace(inp: RefInputs, prec=32, p0=0, p1=None):
    n = inp.width * inp.height
    p1 = n if p1 is None else p1
    out = np.zeros(p1 - p0, RAY_RESULT)
    lib().orc_ref_trace(C.byref(inp.s), C.c_int(prec), C.c_uint64(p0), C.c_uint64(p1), _p(out))
    return out


def ref_eval_object(inp: RefInputs, oi, pts, prec=64):
    """evaluate_object of object oi at world points pts (n x 3):
    (distances float64, materials uint32)."""
    pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
    d = np.zeros(pts.shape[0])
    m = np.zeros(pts.shape[0], np.uint32)
    f = lib().orc_ref_eval_object
    f.argtypes = [C.c_void_p, C.c_int, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]
    for i in range(pts.shape[0]):
        dd = C.c_double(0)
        mm = C.c_uint32(0)
        f(C.byref(inp.s), prec, oi, pts[i].ctypes.data, C.byref(dd), C.byref(mm))
        d[i], m[i] = dd.value, mm.value
    return d, m


def sparse_from_objects(inp: RefInputs, vs, bmin, bmax):
    """orc_sparse_from_objects: (cells, l2, l1, dist, mat, aux) of the
    resampled scene."""
    f = lib().orc_sparse_from_objects
    f.argtypes = [C.c_void_p, C.c_float] + [C.c_void_p] * 10
    lo, hi = np.asarray(bmin, np.float32), np.asarray(bmax, np.float32)
    cells = np.zeros(3, np.uint32)
    nb = np.zeros(1, np.uint32)
    nr = np.zeros(1, np.uint64)
    f(C.byref(inp.s), vs, _p(lo), _p(hi), _p(cells), _p(nb), _p(nr), None, None, None, None, None)
    l2 = np.zeros(int(np.prod(cells.astype(np.int64))), np.uint32)
    l1 = np.zeros(int(nb[0]) * 512, np.uint32)
    dist = np.zeros(int(nr[0]) * 512, np.uint16)
    mat = np.zeros(int(nr[0]) * 512, np.uint8)
    aux = np.zeros(int(nr[0]) * 512, np.uint16)
    f(C.byref(inp.s), vs, _p(lo), _p(hi), _p(cells), _p(nb), _p(nr), _p(l2), _p(l1), _p(dist), _p(mat), _p(aux))
    return cells, l2, l1, dist, mat, aux


def ref_trace_parallel(inp: RefInputs, threads=0):
    out = np.zeros(inp.width * inp.height, RAY_RESULT)
    used = lib().orc_ref_trace_parallel(C.byref(inp.s), _p(out), C.c_int(threads))
    return out, used


class SparseInputs:
    """A sparse scene as host arrays (l2, l1, dist f16 bits, mat u8)."""

    def __init__(self, origin, voxel_size, cells, l2, l1, dist, mat, aux=None):
        self.aux = None if aux is None else np.ascontiguousarray(aux, np.uint16)
        self.l2 = np.ascontiguousarray(l2, np.uint32)
        self.l1 = np.ascontiguousarray(l1, np.uint32)
        self.dist = np.ascontiguousarray(dist, np.uint16)
        self.mat = np.ascontiguousarray(mat, np.uint8)
        cells = [int(c) for c in cells]
        nc = cells[0] * cells[1] * cells[2]
        self.free_l2 = np.zeros(nc, np.uint8)
        self.oct = np.zeros(nc, np.uint32)
        self.bext = np.zeros(nc * 512, np.uint32)
        s = SparseScene()
        s.origin[:] = [float(x) for x in origin]
        s.voxel_size = float(voxel_size)
        s.cells[:] = cells
        s.l2 = _p(self.l2).value
        s.l1 = _p(self.l1).value
        s.dist = _p(self.dist).value
        s.mat = _p(self.mat).value
        s.free_l2 = _p(self.free_l2).value
        s.oct = _p(self.oct).value
        s.bext = _p(self.bext).value
        s.aux = _p(self.aux).value
        self.s = s
        lib().orc_sparse_free_l2(C.byref(s), _p(self.free_l2))
        lib().orc_sparse_oct(C.byref(s), _p(self.free_l2), _p(self.oct))
        lib().orc_sparse_bext(C.byref(s), _p(self.free_l2), _p(self.oct), _p(self.bext))

    def brick_free(self, bx, by, bz):
        return bool(lib().orc_sparse_brick_free(C.byref(self.s), C.c_uint32(bx), C.c_uint32(by), C.c_uint32(bz)))


def sparse_trace(sp: SparseInputs, cam, su, ext, width, height, prec=32, p0=0, p1=None):
    n = width * height
    p1 = n if p1 is None else p1
    out = np.zeros(p1 - p0, RAY_RESULT)
    cam = np.ascontiguousarray(cam, L.CAMERA_UNIFORMS)
    su = np.ascontiguousarray(su, L.SCENE_UNIFORMS)
    ext = np.ascontiguousarray(ext, L.SCENE_EXT)
    lib().orc_sparse_trace(C.byref(sp.s), _p(cam), _p(su), _p(ext), C.c_uint32(width), C.c_uint32(height),
                           C.c_int(prec), C.c_uint64(p0), C.c_uint64(p1), _p(out))
    return out


def sparse_trace_parallel(sp: SparseInputs, cam, su, ext, width, height, p0=0, p1=None, threads=0):
    n = width * height
    p1 = n if p1 is None else p1
    out = np.zeros(p1 - p0, RAY_RESULT)
    cam = np.ascontiguousarray(cam, L.CAMERA_UNIFORMS)
    su = np.ascontiguousarray(su, L.SCENE_UNIFORMS)
    ext = np.ascontiguousarray(ext, L.SCENE_EXT)
    used = lib().orc_sparse_trace_parallel(C.byref(sp.s), _p(cam), _p(su), _p(ext), C.c_uint32(width),
                                           C.c_uint32(height), C.c_uint64(p0), C.c_uint64(p1), _p(out),
                                           C.c_int(threads))
    return out, used


def sparse_tile_counts(sp: SparseInputs, cam, su, width, height):
    """Cells binned into each 8x8 tile (FP32 binning), shape (tiles_y, tiles_x)."""
    tx, ty = (width + 7) // 8, (height + 7) // 8
    out = np.zeros((ty, tx), np.uint32)
    cam = np.ascontiguousarray(cam, L.CAMERA_UNIFORMS)
    su = np.ascontiguousarray(su, L.SCENE_UNIFORMS)
    lib().orc_sparse_tile_counts(C.byref(sp.s), _p(cam), _p(su), C.c_uint32(width), C.c_uint32(height), _p(out))
    return out


def sparse_tiles(sp: SparseInputs, cam, su, width, height, prec=32):
    """(t_lo, t_hi) per 8x8 tile, shape (tiles_y, tiles_x, 2); prec 32 as float32."""
    tx, ty = (width + 7) // 8, (height + 7) // 8
    out = np.zeros((ty, tx, 2), np.float64)
    cam = np.ascontiguousarray(cam, L.CAMERA_UNIFORMS)
    su = np.ascontiguousarray(su, L.SCENE_UNIFORMS)
    lib().orc_sparse_tiles(C.byref(sp.s), _p(cam), _p(su), C.c_uint32(width), C.c_uint32(height), C.c_int(prec),
                           _p(out))
    return out.astype(np.float32) if prec == 32 else out


def sparse_query(sp: SparseInputs, xyz, prec=32):
    xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
    out = np.zeros(xyz.shape[0], L.POINT_RESULT)
    lib().orc_sparse_query(C.byref(sp.s), _p(xyz), C.c_uint64(xyz.shape[0]), C.c_int(prec), _p(out))
    return out


def sparse_cast(sp: SparseInputs, casts, prec=32):
    c = np.ascontiguousarray(casts, L.SPHERE_CAST)
    out = np.zeros(c.size, L.CAST_RESULT)
    lib().orc_sparse_cast(C.byref(sp.s), _p(c), C.c_uint64(c.size), C.c_int(prec), _p(out))
    return out


def sparse_shadow(sp: SparseInputs, origins, light, max_dist=25.0, k=16.0, prec=32):
    o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
    L_ = np.asarray(light, np.float32)
    out = np.zeros(o.shape[0], np.float32)
    lib().orc_sparse_shadow(C.byref(sp.s), _p(o), C.c_uint64(o.shape[0]), _p(L_), C.c_float(max_dist),
                            C.c_float(k), C.c_int(prec), _p(out))
    return out


def sparse_sample(sp: SparseInputs, p, prec=32):
    p = np.asarray(p, np.float32)
    return lib().orc_sparse_sample(C.byref(sp.s), _p(p), C.c_int(prec))


def pack_gbuffer(res):
    n = res.size
    pos = np.zeros((n, 4), np.float32)
    nrm = np.zeros((n, 4), np.uint16)
    mat = np.zeros(n, np.uint32)
    mot = np.zeros((n, 4), np.float32)
    r = np.ascontiguousarray(res)
    lib().orc_pack_gbuffer(_p(r), C.c_uint64(n), _p(pos), _p(nrm), _p(mat), _p(mot))
    return pos, nrm, mat, mot


def sparse_oct(cells, free_l2):
    """orc_sparse_oct on a bare cell grid (only `cells` and the free flags
    matter): per cell 8 octant extents, 4 bits each."""
    cells = [int(c) for c in cells]
    s = SparseScene()
    s.cells[:] = cells
    f = np.ascontiguousarray(free_l2, np.uint8)
    out = np.zeros(f.size, np.uint32)
    lib().orc_sparse_oct(C.byref(s), _p(f), _p(out))
    return out


def sparse_edit(sp, ops, in_place=False):
    """orc_sparse_edit on copies of the scene's planes (or on the scene's own
    planes with in_place): returns (dist, mat, sorted modified slots)."""
    ops = np.ascontiguousarray(ops, L.EDIT)
    dist = sp.dist if in_place else sp.dist.copy()
    mat = sp.mat if in_place else sp.mat.copy()
    n_bricks = dist.size // 512
    mod = np.zeros(max(n_bricks, 1), np.uint8)
    lib().orc_sparse_edit(C.byref(sp.s), _p(dist), _p(mat), _p(ops), C.c_uint32(ops.size), _p(mod))
    return dist, mat, np.flatnonzero(mod[:n_bricks]).astype(np.uint32)


def sparse_contacts(sp, bodies):
    """orc_sparse_contacts: (contacts[n, 16], counts[n])."""
    b = np.ascontiguousarray(bodies, L.BODY)
    out = np.zeros((max(b.size, 1), 16), L.CONTACT)
    counts = np.zeros(max(b.size, 1), np.uint32)
    lib().orc_sparse_contacts(C.byref(sp.s), _p(b), C.c_uint32(b.size), _p(out), _p(counts))
    return out[: b.size], counts[: b.size]

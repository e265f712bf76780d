"""numpy views of the reference's byte layouts (mirrors include/rkf_march.h).

Each dtype is byte-identical to the Rust ``#[repr(C)]`` struct it names:

* ``VOXEL_SAMPLE``   crates/rkf-core/src/voxel.rs:15-20            (8 B)
* ``GPU_OBJECT``     crates/rkf-render/src/gpu_object.rs:41-95     (256 B)
* ``CAMERA_UNIFORMS`` crates/rkf-render/src/camera.rs:194-213      (144 B)
* ``SCENE_UNIFORMS`` crates/rkf-render/src/gpu_scene.rs:21-43      (16 B)
* ``COARSE_FIELD_UNIFORMS`` crates/rkf-render/src/coarse_field.rs:31-42 (48 B)
* ``BVH_NODE``       crates/rkf-render/src/gpu_bvh.rs:24-41        (32 B)
"""
import numpy as np

EMPTY_SLOT = 0xFFFFFFFF
INTERIOR_SLOT = 0xFFFFFFFE
BVH_INVALID = 0xFFFFFFFF
BRICK_DIM = 8
TILE_SIZE = 16
TILE_MAX_OBJECTS = 32
SDF_TYPE_NONE, SDF_TYPE_ANALYTICAL, SDF_TYPE_VOXELIZED = 0, 1, 2
VOX_GEOMETRY_FIRST, VOX_BLEND = 1, 2
PRIM_SPHERE, PRIM_BOX, PRIM_CAPSULE, PRIM_TORUS, PRIM_CYLINDER, PRIM_PLANE = range(6)
EPS_ABSOLUTE, EPS_RELATIVE = 0, 1
STRATEGY_AUTO, STRATEGY_TILED, STRATEGY_BVH, STRATEGY_BRUTE = 0, 1, 2, 3
MARCH_TILE_CULL = 1
MAX_FLOAT = np.float32(3.402823e38)

VOXEL_SAMPLE = np.dtype([("word0", "<u4"), ("word1", "<u4")])

GPU_OBJECT = np.dtype([
    ("inverse_world", "<f4", (16,)),
    ("aabb_min", "<f4", (4,)),
    ("aabb_max", "<f4", (4,)),
    ("brick_map_offset", "<u4"),
    ("brick_map_dims", "<u4", (3,)),
    ("voxel_size", "<f4"),
    ("material_id", "<u4"),
    ("sdf_type", "<u4"),
    ("blend_mode", "<u4"),
    ("blend_radius", "<f4"),
    ("sdf_params", "<f4", (4,)),
    ("accumulated_scale", "<f4", (3,)),
    ("lod_level", "<u4"),
    ("object_id", "<u4"),
    ("primitive_type", "<u4"),
    ("geometry_aabb_min", "<f4", (3,)),
    ("geometry_aabb_max", "<f4", (3,)),
    ("_padding", "<f4", (15,)),
])

CAMERA_UNIFORMS = np.dtype([
    ("position", "<f4", (4,)),
    ("forward", "<f4", (4,)),
    ("right", "<f4", (4,)),
    ("up", "<f4", (4,)),
    ("resolution", "<f4", (2,)),
    ("jitter", "<f4", (2,)),
    ("prev_vp", "<f4", (16,)),
])

SCENE_UNIFORMS = np.dtype([
    ("num_objects", "<u4"),
    ("max_steps", "<u4"),
    ("max_distance", "<f4"),
    ("hit_threshold", "<f4"),
])

COARSE_FIELD_UNIFORMS = np.dtype([
    ("origin_cam_rel", "<f4", (4,)),
    ("dims", "<u4", (4,)),
    ("voxel_size", "<f4"),
    ("inv_voxel_size", "<f4"),
    ("_pad", "<f4", (2,)),
])

BVH_NODE = np.dtype([
    ("aabb_min", "<f4", (3,)),
    ("left", "<u4"),
    ("aabb_max", "<f4", (3,)),
    ("right_or_object", "<u4"),
])

SCENE_EXT = np.dtype([
    ("hit_eps_mode", "<u4"),
    ("strategy", "<u4"),
    ("shadow_enable", "<u4"),
    ("march_flags", "<u4"),
    ("light_dir", "<f4", (3,)),
    ("shadow_max_dist", "<f4"),
    ("shadow_k", "<f4"),
    ("_reserved1", "<f4", (3,)),
])

DEBUG_PX = np.dtype([
    ("normal", "<f4", (3,)),
    ("brick_slot", "<u4"),
    ("voxel_index", "<u4"),
    ("cell_index", "<u4"),
    ("steps", "<u4"),
    ("evals", "<u4"),
    ("shadow_steps", "<u4"),
    ("shadow_evals", "<u4"),
    ("_pad", "<u4", (2,)),
])

PRIM = np.dtype([
    ("centre", "<f4", (3,)),
    ("type", "<u4"),
    ("params", "<f4", (4,)),
    ("rot", "<f4", (9,)),
    ("group", "<u4"),
    ("material", "<u4"),
    ("seed", "<u4"),
])

SPARSE_PARAMS = np.dtype([
    ("voxel_size", "<f4"),
    ("smin_k", "<f4"),
    ("mat_seed", "<u8"),
    ("bounds_min", "<f4", (3,)),
    ("bounds_max", "<f4", (3,)),
    ("max_bricks", "<u8"),
])

SPARSE_INFO = np.dtype([
    ("origin", "<f4", (3,)),
    ("voxel_size", "<f4"),
    ("cells", "<u4", (3,)),
    ("n_blocks", "<u4"),
    ("n_bricks", "<u8"),
    ("device_bytes", "<u8"),
])

POINT_RESULT = np.dtype([
    ("distance", "<f4"),
    ("material", "<u4"),
    ("inside", "<u4"),
    ("gradient", "<f4", (3,)),
])
SPHERE_CAST = np.dtype([("origin", "<f4", (3,)), ("radius", "<f4"), ("dir", "<f4", (3,)), ("max_t", "<f4")])
CAST_RESULT = np.dtype([("t", "<f4"), ("hit", "<u4"), ("steps", "<u4"), ("normal", "<f4", (3,)),
                        ("distance", "<f4"), ("_pad", "<u4")])
assert SPHERE_CAST.itemsize == 32 and CAST_RESULT.itemsize == 32

GBUFFER_DESC = np.dtype([("position", "<u8"), ("normal", "<u8"), ("material", "<u8"), ("motion", "<u8"),
                         ("debug", "<u8"), ("shadow", "<u8"), ("width", "<u4"), ("height", "<u4")])

BATCH_VIEW, BATCH_SHADOW, BATCH_POINTS, BATCH_SPHERES = 0, 1, 2, 3
SCHEDULE_GRID, SCHEDULE_PERSISTENT, SCHEDULE_GRID_LOWLAT = 0, 1, 2

BATCH_ITEM = np.dtype([
    ("kind", "<u4"), ("n", "<u4"), ("scene", "<u8"),
    ("camera", CAMERA_UNIFORMS), ("su", SCENE_UNIFORMS), ("ext", SCENE_EXT), ("out", GBUFFER_DESC),
    ("origins", "<u8"), ("visibility", "<u8"), ("xyz", "<u8"), ("results", "<u8"),
    ("casts", "<u8"), ("cast_results", "<u8"),
])
assert BATCH_ITEM.itemsize == 328

BATCH_TUNING = np.dtype([("schedule", "<u4"), ("tile_w", "<u4"), ("tile_h", "<u4"), ("waves_per_cu", "<u4")])

(EDIT_CSG_UNION, EDIT_CSG_SUBTRACT, EDIT_CSG_INTERSECT, EDIT_SMOOTH_UNION, EDIT_SMOOTH_SUBTRACT, EDIT_SMOOTH,
 EDIT_FLATTEN, EDIT_PAINT, EDIT_COLOR_PAINT) = range(9)
SHAPE_SPHERE, SHAPE_BOX, SHAPE_CAPSULE, SHAPE_CYLINDER, SHAPE_TORUS, SHAPE_PLANE = range(6)
FALLOFF_LINEAR, FALLOFF_SMOOTH, FALLOFF_SHARP = range(3)

EDIT = np.dtype([
    ("position", "<f4", (3,)), ("edit_type", "<u4"), ("rotation", "<f4", (4,)), ("dimensions", "<f4", (3,)),
    ("shape_type", "<u4"), ("strength", "<f4"), ("blend_k", "<f4"), ("falloff", "<u4"), ("material_id", "<u4"),
])

RKF_HEADER = np.dtype([("magic", "S4"), ("version", "<u4"), ("lod_count", "<u4"), ("material_count", "<u4"),
                       ("aabb_min", "<f4", (3,)), ("aabb_max", "<f4", (3,)), ("analytical_type", "<u4"),
                       ("analytical_params", "<f4", (4,)), ("material_ids", "<u2", (16,)), ("_reserved", "u1", (36,))])
RKF_LOD = np.dtype([("voxel_size", "<f4"), ("brick_count", "<u4"), ("brick_dims", "<u4", (3,)), ("_pad", "<u4"),
                    ("data_offset", "<u8"), ("compressed_size", "<u4"), ("uncompressed_size", "<u4")])
RKF_LOD_SRC = np.dtype([("voxel_size", "<f4"), ("dims", "<u4", (3,)), ("map", "<u8"), ("bricks", "<u8"),
                        ("brick_count", "<u4"), ("_pad", "<u4")])

BODY_SPHERE, BODY_BOX, BODY_CAPSULE = 0, 1, 2
MAX_CONTACTS = 16
BODY = np.dtype([("position", "<f4", (3,)), ("shape", "<u4"), ("rotation", "<f4", (4,)), ("dims", "<f4", (3,)),
                 ("threshold", "<f4")])
CONTACT = np.dtype([("position", "<f4", (3,)), ("normal", "<f4", (3,)), ("penetration", "<f4"), ("sample", "<u4")])

SERVE_BATCHED, SERVE_PER_REQUEST = 0, 1
REQ_KEEP_ON_DEVICE = 1
REQ_QUEUED, REQ_ADMITTED, REQ_RUNNING, REQ_DONE, REQ_FAILED = 0, 1, 2, 3, 4

SERVER_CONFIG = np.dtype([
    ("memory_budget", "<u8"), ("max_requests", "<u4"), ("rounds_in_flight", "<u4"),
    ("max_round_work", "<u8"), ("max_round_requests", "<u4"), ("policy", "<u4"),
    ("tuning", BATCH_TUNING), ("max_pending", "<u4"), ("_reserved", "<u4"),
])

REQUEST = np.dtype([
    ("kind", "<u4"), ("scene", "<u4"), ("n", "<u4"), ("flags", "<u4"), ("width", "<u4"), ("height", "<u4"),
    ("camera", CAMERA_UNIFORMS), ("su", SCENE_UNIFORMS), ("ext", SCENE_EXT),
    ("in", "<u8"), ("out", "<u8"), ("view_out", GBUFFER_DESC), ("tag", "<u8"),
])

REQUEST_STATUS = np.dtype([
    ("state", "<u4"), ("error", "<i4"), ("tag", "<u8"), ("round", "<u8"), ("round_requests", "<u4"),
    ("_pad", "<u4"), ("t_submit", "<f8"), ("t_admit", "<f8"), ("t_launch", "<f8"), ("t_done", "<f8"),
])

SERVER_STATS = np.dtype([
    ("submitted", "<u8"), ("completed", "<u8"), ("failed", "<u8"), ("rounds", "<u8"), ("work_items", "<u8"),
    ("arena_bytes", "<u8"), ("arena_in_use", "<u8"), ("arena_peak", "<u8"),
    ("max_round_requests_seen", "<u4"), ("queued", "<u4"), ("admitted", "<u4"), ("running", "<u4"),
    ("gpu_seconds", "<f8"),
])

assert VOXEL_SAMPLE.itemsize == 8
assert GPU_OBJECT.itemsize == 256
assert CAMERA_UNIFORMS.itemsize == 144
assert SCENE_UNIFORMS.itemsize == 16
assert COARSE_FIELD_UNIFORMS.itemsize == 48
assert BVH_NODE.itemsize == 32
assert SCENE_EXT.itemsize == 48
assert DEBUG_PX.itemsize == 48
assert PRIM.itemsize == 80
assert POINT_RESULT.itemsize == 24
assert GBUFFER_DESC.itemsize == 56
assert BATCH_ITEM.itemsize == 8 + 8 + 144 + 16 + 48 + 56 + 48
assert EDIT.itemsize == 64
assert RKF_HEADER.itemsize == 128 and RKF_LOD.itemsize == 40 and RKF_LOD_SRC.itemsize == 40
assert BODY.itemsize == 48
assert CONTACT.itemsize == 32
assert SERVER_CONFIG.itemsize == 56
assert REQUEST.itemsize == 312
assert REQUEST_STATUS.itemsize == 64
assert SERVER_STATS.itemsize == 88


def scene_uniforms(num_objects=0, max_steps=128, max_distance=100.0, hit_threshold=0.001):
    """SceneUniforms::default (gpu_scene.rs:34-43) with overrides."""
    s = np.zeros((), SCENE_UNIFORMS)
    s["num_objects"] = num_objects
    s["max_steps"] = max_steps
    s["max_distance"] = max_distance
    s["hit_threshold"] = hit_threshold
    return s


def scene_ext(hit_eps_mode=EPS_ABSOLUTE, strategy=STRATEGY_AUTO, shadow_enable=0,
              light_dir=(0.0, 1.0, 0.0), shadow_max_dist=25.0, shadow_k=16.0, march_flags=0):
    e = np.zeros((), SCENE_EXT)
    e["hit_eps_mode"] = hit_eps_mode
    e["strategy"] = strategy
    e["shadow_enable"] = shadow_enable
    e["march_flags"] = march_flags
    e["light_dir"] = light_dir
    e["shadow_max_dist"] = shadow_max_dist
    e["shadow_k"] = shadow_k
    return e


def edit_op(position, edit_type, shape_type, dimensions, strength=1.0, blend_k=0.0, falloff=FALLOFF_SMOOTH,
            material_id=0, rotation=(0.0, 0.0, 0.0, 1.0)):
    """An EditOp record (rkf-edit edit_op.rs:15-28)."""
    e = np.zeros((), EDIT)
    e["position"] = position
    e["edit_type"] = edit_type
    e["rotation"] = rotation
    e["dimensions"] = dimensions
    e["shape_type"] = shape_type
    e["strength"] = strength
    e["blend_k"] = blend_k
    e["falloff"] = falloff
    e["material_id"] = material_id
    return e

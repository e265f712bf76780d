"""rkifield_amd — MI355X (gfx950) SDF ray-march hot path of joeleaver/rkifield.

The compute lives in librkf_march.so (hand-written HIP kernels behind the C ABI
of include/rkf_march.h); this package is its Python mirror. Importing it never
falls back to a CPU path: a missing library raises when first used.
"""
from . import layouts  # noqa: F401
from .march import (  # noqa: F401
    Batch,
    BatchGraph,
    CoarseField,
    Context,
    DeviceArray,
    GBuffer,
    GpuScene,
    PinnedArray,
    RayMarchPass,
    RkfMarchError,
    Server,
    SparseScene,
    TileObjectCullPass,
    build_bvh,
    camera_look_at,
    lib,
    voxelize_sphere,
    voxelize_test_cross,
)

__version__ = "0.1.0"

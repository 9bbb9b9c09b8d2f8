"""MI355X (gfx950) sparse-voxel-octree ray traversal, block cache and feedback pipeline.

The compute path is librayt_hip.so (hand-written HIP kernels + C ABI,
include/rayt_hip.h). This package is its Python host binding and the
renderer-facing mirror of the reference API (rayt_amd.api).
"""
from ._lib import LIB_PATH, RaytError, lib  # noqa: F401

__all__ = ["LIB_PATH", "RaytError", "lib"]

"""Core module exports (same names as atom3d.core)."""
from .data_structures import (
    AABBIntersectResult,
    ClosestPointResult,
    RayIntersectResult,
    SegmentIntersectResult,
    TriangleIntersectResult,
    VisibilityResult,
    VoxelFaceMapping,
    VoxelPolygonMapping,
)
from .device_utils import ensure_same_device, get_default_cuda_device, resolve_device
from .mesh_bvh import MeshBVH

__all__ = [
    "MeshBVH", "AABBIntersectResult", "RayIntersectResult", "SegmentIntersectResult", "ClosestPointResult",
    "TriangleIntersectResult", "VoxelFaceMapping", "VoxelPolygonMapping", "VisibilityResult",
    "resolve_device", "ensure_same_device", "get_default_cuda_device",
]

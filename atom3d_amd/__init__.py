"""atom3d_amd: MI355X-native (gfx950 HIP) implementation of Atom3D's geometry
hot path, with the reference's Python API (``import atom3d_amd as atom3d``).

GPU queries run in the hand-written HIP library ``_lib/libatom3d_amd.so``
(build with ``python -m atom3d_amd.build``); there is no CPU fallback.
"""
__version__ = "0.1.0"

from .core.data_structures import (
    AABBIntersectResult,
    ClosestPointResult,
    RayIntersectResult,
    SegmentIntersectResult,
    TriangleIntersectResult,
    VisibilityResult,
    VoxelFaceMapping,
    VoxelPolygonMapping,
)
from .core.mesh_bvh import MeshBVH
from .grid.cube_grid import CubeGrid
from .grid.octree_indexer import OctreeIndexer

__all__ = [
    "MeshBVH", "AABBIntersectResult", "RayIntersectResult", "SegmentIntersectResult", "ClosestPointResult",
    "TriangleIntersectResult", "VoxelFaceMapping", "VoxelPolygonMapping", "VisibilityResult",
    "OctreeIndexer", "CubeGrid",
]

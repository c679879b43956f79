"""Result types of the drop-in API.

Field names, order and defaults match Atom3D atom3d/core/data_structures.py:10-108
so code written against the reference reads the same attributes.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import torch


@dataclass
class AABBIntersectResult:
    """Box-mesh intersection: hit [N] bool; pairs (mode >= 1); clip data (mode >= 2)."""
    hit: torch.Tensor
    aabb_ids: Optional[torch.Tensor] = None
    face_ids: Optional[torch.Tensor] = None
    centroids: Optional[torch.Tensor] = None
    areas: Optional[torch.Tensor] = None
    poly_verts: Optional[torch.Tensor] = None
    poly_counts: Optional[torch.Tensor] = None


@dataclass
class RayIntersectResult:
    hit: torch.Tensor
    t: torch.Tensor
    face_ids: torch.Tensor
    hit_points: torch.Tensor
    normals: torch.Tensor
    bary_coords: Optional[torch.Tensor] = None


@dataclass
class RayAllHitResult:
    ray_idx: torch.Tensor
    t: torch.Tensor
    face_ids: torch.Tensor
    sign: torch.Tensor
    hit_points: torch.Tensor
    bary_coords: torch.Tensor


@dataclass
class SegmentIntersectResult:
    hit: torch.Tensor
    hit_points: torch.Tensor
    face_ids: torch.Tensor
    bary_coords: torch.Tensor
    segment_ids: Optional[torch.Tensor] = None


@dataclass
class ClosestPointResult:
    """Closest-point (UDF/SDF) query: distances [N], face_ids [N] int32,
    closest_points [N,3], uvw [N,3]."""
    distances: torch.Tensor
    face_ids: torch.Tensor
    closest_points: torch.Tensor
    uvw: Optional[torch.Tensor] = None
    pseudo_normals: Optional[torch.Tensor] = None


@dataclass
class TriangleIntersectResult:
    edge_hit: torch.Tensor
    hit_points: torch.Tensor
    hit_face_ids: torch.Tensor
    hit_edge_ids: torch.Tensor


@dataclass
class VoxelFaceMapping:
    """Voxel -> faces in CSR form."""
    voxel_coords: torch.Tensor
    face_indices: torch.Tensor
    face_start: torch.Tensor
    face_count: torch.Tensor

    def get_faces_for_voxel(self, voxel_idx: int) -> torch.Tensor:
        s = int(self.face_start[voxel_idx])
        return self.face_indices[s:s + int(self.face_count[voxel_idx])]


@dataclass
class VoxelPolygonMapping:
    voxel_coords: torch.Tensor
    polygons: torch.Tensor
    polygon_counts: torch.Tensor
    face_indices: torch.Tensor
    voxel_ids: torch.Tensor

    def get_polygon(self, idx: int) -> torch.Tensor:
        return self.polygons[idx, :int(self.polygon_counts[idx])]


@dataclass
class VisibilityResult:
    visibility: torch.Tensor
    visible_mask: Optional[torch.Tensor] = None
    hit_distances: Optional[torch.Tensor] = None

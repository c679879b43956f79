"""Voxelizer: surface voxelisation of a mesh on a CubeGrid, on the GPU.

Same class, methods and results as Atom3D atom3d/apps/voxelizer.py:15-204.
The reference materialises every cell of the block covering the faces'
boxes (or of the whole grid), forms their boxes and runs
MeshBVH.intersect_aabb on all of them.  Here one HIP kernel enumerates the
block implicitly and keeps the cells some face intersects (first-hit SAT
walk, order-preserving compaction); only those cells get the full pair walk
(count, exclusive sum, write) and a radix sort into (cell, face) order.  The
cell set, its order and the cell boxes are the reference's; within a voxel
the faces come in ascending id (the reference's order there follows its
atomics and an unstable argsort).
"""
from __future__ import annotations

import torch

from ..core.data_structures import VoxelFaceMapping, VoxelPolygonMapping
from ..core.mesh_bvh import MeshBVH
from ..grid.cube_grid import CubeGrid


class Voxelizer:
    """Mesh voxelisation (voxelizer.py:15-34): a MeshBVH and a CubeGrid."""

    def __init__(self, bvh: MeshBVH, grid: CubeGrid):
        self.bvh = bvh
        self.grid = grid

    def _geom(self):
        return self.grid.bounds[0].tolist(), self.grid.cell_size.tolist()

    def _block(self, strategy: str):
        if strategy == "all":
            return [0, 0, 0], [self.grid.res] * 3
        lo, hi = self.bvh.get_face_aabb()
        return self.grid.candidate_block(lo, hi)

    def voxelize_surface(self, strategy: str = "candidate") -> torch.Tensor:
        """int64 [K,3] cells some face intersects (SAT), meshgrid('ij') order
        of the candidate block ('candidate') or of the whole grid ('all')
        (voxelizer.py:36-73)."""
        origin, dims = self._block(strategy)
        b0, cell = self._geom()
        return self.bvh._bvh.block_cells(b0, cell, origin, dims)

    def voxelize_with_faces(self) -> VoxelFaceMapping:
        """CSR voxel -> faces map of the candidate block (voxelizer.py:75-131):
        voxel_coords int64 [K,3], face_indices / face_start / face_count
        int32."""
        origin, dims = self._block("candidate")
        b0, cell = self._geom()
        vox = self.bvh._bvh.block_cells(b0, cell, origin, dims)
        counts, _, faces = self.bvh._bvh.cells_sat_pairs(vox, b0, cell)
        start = torch.cumsum(counts, 0, dtype=torch.int32) - counts
        return VoxelFaceMapping(voxel_coords=vox, face_indices=faces, face_start=start, face_count=counts)

    def voxelize_with_polygons(self, max_polygon_verts: int = 8) -> VoxelPolygonMapping:
        """voxelize_with_faces plus, per (voxel, face) pair, the face's three
        vertices as its polygon (count 3), as the reference does
        (voxelizer.py:133-204); the exact clipped polygons are
        MeshBVH.intersect_aabb(mode=3)."""
        vf = self.voxelize_with_faces()
        dev = self.bvh.device
        P = int(vf.face_indices.shape[0])
        polygons = torch.zeros((P, max_polygon_verts, 3), dtype=torch.float32, device=dev)
        if P == 0:
            return VoxelPolygonMapping(voxel_coords=vf.voxel_coords, polygons=polygons,
                                       polygon_counts=torch.empty(0, dtype=torch.int32, device=dev),
                                       face_indices=torch.empty(0, dtype=torch.int32, device=dev),
                                       voxel_ids=torch.empty(0, dtype=torch.int32, device=dev))
        polygons[:, :3] = self.bvh.vertices[self.bvh.faces[vf.face_indices.long()].long()]
        voxel_ids = torch.repeat_interleave(torch.arange(vf.voxel_coords.shape[0], dtype=torch.int32, device=dev),
                                            vf.face_count.long())
        return VoxelPolygonMapping(voxel_coords=vf.voxel_coords, polygons=polygons,
                                   polygon_counts=torch.full((P,), 3, dtype=torch.int32, device=dev),
                                   face_indices=vf.face_indices, voxel_ids=voxel_ids)


__all__ = ["Voxelizer"]

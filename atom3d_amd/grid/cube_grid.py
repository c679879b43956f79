"""CubeGrid: regular cube grid over an axis-aligned box.

Mirrors the parts of Atom3D atom3d/grid/cube_grid.py the hot path uses:
constructor and attributes (:48-114), coordinate conversions (:158-169),
ravel/unravel (:171-184), cube indexing, boxes and corners (:190-235) and
voxel_unique_vertices (:237-260).  Cube coordinates are int64 (``index_dtype``),
positions float32; boxes are ``ijk.float() * cell + b0`` with two separate
float32 roundings, exactly like the reference.  voxel_unique_vertices runs
on the GPU (one radix sort of the corner keys, ranks by scan) and needs
device tensors and cubes inside the grid.
"""
from __future__ import annotations

from typing import Optional, Tuple, Union

import torch

from ..constants import CUBE_CORNERS, CUBE_EDGES, CUBE_FACES, EDGE_TO_FACES
from ..core.device_utils import resolve_device

FACE_NORMALS = torch.tensor([[0, 0, -1], [0, 0, 1], [-1, 0, 0], [1, 0, 0], [0, -1, 0], [0, 1, 0]], dtype=torch.int64)


class CubeGrid:
    def __init__(self, resolution: int, bounds: Optional[torch.Tensor] = None,
                 device: Optional[Union[str, torch.device]] = None):
        if not isinstance(resolution, int) or resolution < 1:
            raise ValueError(f"resolution must be positive int, got {resolution}")
        self.res = int(resolution)
        self.device = resolve_device(bounds, device=device, default="cuda")
        self.dtype = torch.float32
        self.index_dtype = torch.int64
        self.num_vertices_per_axis = self.res + 1
        self.num_cubes = self.res ** 3
        self.num_vertices = self.num_vertices_per_axis ** 3
        if bounds is not None:
            if tuple(bounds.shape) != (2, 3):
                raise ValueError(f"bounds must be [2,3], got {tuple(bounds.shape)}")
            self.bounds = bounds.to(device=self.device, dtype=self.dtype)
        else:
            self.bounds = torch.tensor([[-1.0] * 3, [1.0] * 3], device=self.device, dtype=self.dtype)
        self.cell_size = (self.bounds[1] - self.bounds[0]) / float(self.res)
        self.coords_1d = torch.linspace(float(self.bounds[0, 0]), float(self.bounds[1, 0]),
                                        self.num_vertices_per_axis, device=self.device, dtype=self.dtype)
        self.CUBE_CORNERS = CUBE_CORNERS.to(self.device)
        self.CUBE_EDGES = CUBE_EDGES.to(self.device)
        self.CUBE_FACES = CUBE_FACES.to(self.device)
        self.EDGE_TO_FACES = EDGE_TO_FACES.to(self.device)
        self.FACE_NORMALS = FACE_NORMALS.to(self.device)
        r, p = self.res, self.num_vertices_per_axis
        self.num_edges = 3 * r * p * p
        self.num_faces = 3 * r * r * p

    # coordinates
    def world_to_grid(self, world_coords: torch.Tensor) -> torch.Tensor:
        return (world_coords.to(self.device, dtype=self.dtype) - self.bounds[0]) / self.cell_size

    def grid_to_world(self, grid_coords: torch.Tensor) -> torch.Tensor:
        return (grid_coords.to(self.device).float() + 0.5) * self.cell_size + self.bounds[0]

    def vertex_to_world(self, vertex_ijk: torch.Tensor) -> torch.Tensor:
        return vertex_ijk.to(self.device).float() * self.cell_size + self.bounds[0]

    # flat indices (row-major, k fastest)
    def ravel_ijk(self, ijk: torch.Tensor, dims: Tuple[int, int, int]) -> torch.Tensor:
        _, ny, nz = dims
        ijk = ijk.to(device=self.device, dtype=self.index_dtype)
        return (ijk[..., 0] * ny + ijk[..., 1]) * nz + ijk[..., 2]

    def unravel_idx(self, idx: torch.Tensor, dims: Tuple[int, int, int]) -> torch.Tensor:
        _, ny, nz = dims
        idx = idx.to(device=self.device, dtype=self.index_dtype)
        k = idx % nz
        j = (idx // nz) % ny
        i = idx // (ny * nz)
        return torch.stack([i, j, k], dim=-1)

    # cubes
    def all_cube_indices(self) -> torch.Tensor:
        return torch.arange(self.num_cubes, device=self.device, dtype=self.index_dtype)

    def cube_to_ijk(self, cube_idx: torch.Tensor) -> torch.Tensor:
        return self.unravel_idx(cube_idx, (self.res,) * 3)

    def ijk_to_cube(self, ijk: torch.Tensor) -> torch.Tensor:
        return self.ravel_ijk(ijk, (self.res,) * 3)

    def cube_aabb(self, cube_idx: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
        lo = self.vertex_to_world(self.cube_to_ijk(cube_idx))
        return lo, lo + self.cell_size

    def cube_center(self, cube_idx: torch.Tensor) -> torch.Tensor:
        return self.grid_to_world(self.cube_to_ijk(cube_idx))

    def cube_corner_vertex_indices(self, cube_idx: torch.Tensor) -> torch.Tensor:
        corners = self.cube_to_ijk(cube_idx)[:, None, :] + self.CUBE_CORNERS[None]
        p = self.num_vertices_per_axis
        return self.ravel_ijk(corners.reshape(-1, 3), (p, p, p)).view(-1, 8)

    def cube_corner_coords(self, cube_idx: torch.Tensor) -> torch.Tensor:
        return self.vertex_to_world(self.cube_to_ijk(cube_idx)[:, None, :] + self.CUBE_CORNERS[None])

    def vertex_ijk_from_indices(self, vertex_idx: torch.Tensor) -> torch.Tensor:
        p = self.num_vertices_per_axis
        return self.unravel_idx(vertex_idx.flatten(), (p, p, p)).view(*vertex_idx.shape, 3)

    def vertex_coords_from_indices(self, vertex_idx: torch.Tensor) -> torch.Tensor:
        return self.vertex_to_world(self.vertex_ijk_from_indices(vertex_idx))

    def generate_all_cells(self) -> torch.Tensor:
        """Every cell's ijk, meshgrid('ij') order (cube_grid.py:525-528)."""
        x = torch.arange(self.res, device=self.device, dtype=self.index_dtype)
        xx, yy, zz = torch.meshgrid(x, x, x, indexing="ij")
        return torch.stack([xx.flatten(), yy.flatten(), zz.flatten()], dim=1)

    def candidate_block(self, aabb_min: torch.Tensor, aabb_max: torch.Tensor):
        """(origin, dims) of the cell block covering boxes [K,3]: floor / ceil
        of their grid coordinates, clamped to the grid, min / max over the
        boxes (cube_grid.py:542-550)."""
        aabb_min = aabb_min.to(self.device, dtype=self.dtype)
        aabb_max = aabb_max.to(self.device, dtype=self.dtype)
        gmin = self.world_to_grid(aabb_min).floor().to(self.index_dtype).clamp(0, self.res - 1)
        gmax = self.world_to_grid(aabb_max).ceil().to(self.index_dtype).clamp(0, self.res - 1)
        lo = gmin.min(dim=0)[0].tolist()
        hi = gmax.max(dim=0)[0].tolist()
        return lo, [h - l + 1 for l, h in zip(lo, hi)]

    def generate_candidate_cells_from_aabb(self, aabb_min: torch.Tensor, aabb_max: torch.Tensor) -> torch.Tensor:
        """All cells of the block covering the boxes, meshgrid('ij') order
        (cube_grid.py:542-557)."""
        lo, dims = self.candidate_block(aabb_min, aabb_max)
        ax = [torch.arange(l, l + d, device=self.device, dtype=self.index_dtype) for l, d in zip(lo, dims)]
        xx, yy, zz = torch.meshgrid(*ax, indexing="ij")
        return torch.stack([xx.flatten(), yy.flatten(), zz.flatten()], dim=1)

    def voxel_unique_vertices(self, voxels: torch.Tensor):
        """(unique vertex indices [M] int64, their coordinates [M,3] float32,
        voxel -> vertex [N,8] int64), vertices in ascending index order."""
        from ..kernels.cells import cell_vertices

        if voxels.dim() == 2 and voxels.shape[1] == 3:
            ijk = voxels
        elif voxels.dim() == 1:
            ijk = self.cube_to_ijk(voxels)
        else:
            raise ValueError(f"voxels must be [N,3] or [N], got {tuple(voxels.shape)}")
        lo = self.bounds[0].tolist()
        cell = self.cell_size.tolist()
        inv, keys, _, world = cell_vertices(ijk.to(self.device), self.res, lo=lo, cell=cell, want_keys=True)
        return keys, world, inv

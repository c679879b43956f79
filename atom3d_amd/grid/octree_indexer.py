"""OctreeIndexer: multi-level cube grid with GPU coarse-to-fine voxelisation.

Mirrors Atom3D atom3d/grid/octree_indexer.py: per-level sizes (:34-75),
level conversions and boxes (:79-128, :243-337), subdivide/merge (:191-240),
all_cubes_at_level (:426-439) and octree_traverse (:470-509).

``octree_traverse`` keeps the reference's semantics and output order (start
from every cube of ``min_level`` in meshgrid('ij') order, keep the cubes some
face intersects, replace each kept cube by its 8 children in CUBE_CORNERS
order, repeat down to ``max_level``) but runs each level as one fused GPU step:
candidate boxes are formed in registers, tested with a first-hit BVH/SAT walk
and compacted in order on the device; only the kept count crosses to the host
per level.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Tuple, Union

import torch

from ..kernels import _lib
from .cube_grid import CubeGrid


class OctreeIndexer(CubeGrid):
    def __init__(self, max_level: int, bounds: Optional[torch.Tensor] = None,
                 device: Optional[Union[str, torch.device]] = None):
        self.max_level = max_level
        self._max_resolution = 2 ** max_level
        super().__init__(resolution=self._max_resolution, bounds=bounds, device=device)
        self._level_resolutions = [2 ** level for level in range(max_level + 1)]
        self._level_cell_sizes = [(self.bounds[1] - self.bounds[0]) / r for r in self._level_resolutions]

    # sizes
    def get_resolution(self, level: int) -> int:
        return self._level_resolutions[level] if level < len(self._level_resolutions) else 2 ** level

    def get_cell_size(self, level: int) -> torch.Tensor:
        if level < len(self._level_cell_sizes):
            return self._level_cell_sizes[level]
        return (self.bounds[1] - self.bounds[0]) / (2 ** level)

    def get_num_cubes(self, level: int) -> int:
        return self.get_resolution(level) ** 3

    # conversions
    def world_to_grid_level(self, world_coords: torch.Tensor, level: int) -> torch.Tensor:
        return ((world_coords - self.bounds[0]) / self.get_cell_size(level)).floor().to(self.index_dtype)

    def grid_to_world_level(self, grid_coords: torch.Tensor, level: int) -> torch.Tensor:
        return (grid_coords.float() + 0.5) * self.get_cell_size(level) + self.bounds[0]

    def get_cell_aabb_level(self, grid_coords: torch.Tensor, level: int) -> Tuple[torch.Tensor, torch.Tensor]:
        cs = self.get_cell_size(level)
        return grid_coords.float() * cs + self.bounds[0], (grid_coords.float() + 1) * cs + self.bounds[0]

    # level transitions
    def subdivide(self, parent_coords: torch.Tensor, parent_level: int) -> torch.Tensor:
        kids = (parent_coords * 2)[:, None, :] + self.CUBE_CORNERS[None]
        return kids.reshape(-1, 3).to(self.index_dtype)

    def merge(self, child_coords: torch.Tensor, child_level: int) -> torch.Tensor:
        return torch.div(child_coords, 2, rounding_mode="floor")

    def get_root_coords(self) -> torch.Tensor:
        return torch.zeros(1, 3, dtype=self.index_dtype, device=self.device)

    def cube_corner_coords_level(self, cube_ijk: torch.Tensor, level: int) -> torch.Tensor:
        corners = cube_ijk[:, None, :] + self.CUBE_CORNERS[None]
        return corners.float() * self.get_cell_size(level) + self.bounds[0]

    def cube_to_ijk_level(self, cube_idx: torch.Tensor, level: Optional[int] = None) -> torch.Tensor:
        r = self.get_resolution(self.max_level if level is None else level)
        return self.unravel_idx(cube_idx, (r, r, r))

    def ijk_to_cube_level(self, ijk: torch.Tensor, level: Optional[int] = None) -> torch.Tensor:
        r = self.get_resolution(self.max_level if level is None else level)
        return self.ravel_ijk(ijk, (r, r, r))

    def cube_aabb_level(self, cubes: torch.Tensor, level: Optional[int] = None) -> Tuple[torch.Tensor, torch.Tensor]:
        level = self.max_level if level is None else level
        if cubes.dim() == 2 and cubes.shape[1] == 3:
            ijk = cubes
        elif cubes.dim() == 1:
            ijk = self.cube_to_ijk_level(cubes, level)
        else:
            raise ValueError(f"cubes must be [B, 3] ijk or [B] linear idx, got shape {tuple(cubes.shape)}")
        cs = self.get_cell_size(level)
        lo = ijk.float() * cs + self.bounds[0]
        return lo, lo + cs

    def all_cubes_at_level(self, level: int) -> torch.Tensor:
        r = self.get_resolution(level)
        x = torch.arange(r, device=self.device, dtype=self.index_dtype)
        g = torch.meshgrid(x, x, x, indexing="ij")
        return torch.stack([a.flatten() for a in g], dim=1)

    # traversal
    def octree_traverse(self, bvh, min_level: int = 2) -> torch.Tensor:
        """Active cubes [K, 3] int64 at ``max_level``, in the reference's order."""
        blob = bvh._bvh.box_blob if hasattr(bvh, "_bvh") else bvh.box_blob  # the reference-layout SAH tree
        dev = blob.device
        b = self.bounds.detach().to("cpu", torch.float32).reshape(-1).tolist()
        bounds = (ctypes.c_float * 6)(*b)
        s = _lib.stream(dev)
        with torch.cuda.device(dev):
            n_kept = torch.empty(1, dtype=torch.int64, device=dev)
            parents = None
            n_par = 0
            self.last_level_counts = []  # (level, cubes tested, cubes hit)
            for level in range(min_level, max(self.max_level, min_level) + 1):
                if level > min_level and n_par == 0:
                    break
                n_cand = (2 ** level) ** 3 if parents is None else 8 * n_par
                wsb = ctypes.c_int64()
                _lib.call("a3d_octree_ws_bytes", n_cand, ctypes.byref(wsb))
                ws = torch.empty(wsb.value, dtype=torch.uint8, device=dev)
                kept = torch.empty(max(n_cand, 1), dtype=torch.int64, device=dev)
                _lib.call("a3d_octree_level", _lib.ptr(blob), bounds, level, _lib.ptr(parents), n_par,
                          _lib.ptr(kept), _lib.ptr(n_kept), _lib.ptr(ws), wsb.value, s)
                n_par = int(n_kept.item())
                self.last_level_counts.append((level, n_cand, n_par))
                parents = kept
                if level >= self.max_level:
                    break
            out = torch.empty(n_par, 3, dtype=torch.int64, device=dev)
            if n_par > 0:
                _lib.call("a3d_octree_unpack", _lib.ptr(parents), n_par, _lib.ptr(out), s)
        return out

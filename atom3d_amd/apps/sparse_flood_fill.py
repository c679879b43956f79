"""sparse_flood_fill: octree voxelisation, flood fill and the dam's k-ring
neighbourhood signs, all on the GPU.

Same function, arguments and returned dict as Atom3D
atom3d/apps/sparse_flood_fill.py:103-166 (with get_dam_neighborhood_signs
:34-100), plus the glue towards DMC: get_dam_wet_faces / _edges / _corners
(:169-227) and construct_surface_dual_graph (:230-291), whose unique / sort /
search steps run as HIP kernels (kernels/cells.py).  The neighbourhood and its signs are computed on the flood's bit
planes (k 26-dilations of the dam plane, then masks) instead of unique/sort/
searchsorted over coordinate lists; the coordinate lists come out in the same
lexicographic order.
"""
from __future__ import annotations

import logging
import time
from typing import Dict, Optional, Tuple

import numpy as np
import torch

from ..constants import NEIGHBOR_OFFSETS_26, SIGN_DRY, SIGN_WATER
from ..core.mesh_bvh import MeshBVH
from ..grid.octree_indexer import OctreeIndexer
from ..kernels.cells import cell_neighbors, cell_vertices, unique_cells, wet_mask
from ..kernels.flood_fill import _out, flood_planes

logger = logging.getLogger(__name__)

_PADDING = 10


def get_dam_neighborhood_signs(dam_ijk: torch.Tensor, water_ijk: torch.Tensor, dry_ijk: torch.Tensor,
                               resolution: int, k_ring: int = 1) -> Tuple[torch.Tensor, torch.Tensor]:
    """k-ring 26-neighbourhood of the dam voxels (unique, lexicographic, int32)
    and its signs (int8): 0 dam, 1 water, -1 dry, 1 otherwise."""
    device = dam_ijk.device
    if len(dam_ijk) == 0:
        return torch.empty(0, 3, dtype=torch.int32, device=device), torch.empty(0, dtype=torch.int8, device=device)
    R = resolution
    offs = NEIGHBOR_OFFSETS_26.to(device).long()
    nb = dam_ijk.long()
    front = nb
    for ring in range(k_ring):
        cand = (front[:, None, :] + offs[None]).reshape(-1, 3)
        cand = cand[((cand >= 0) & (cand < R)).all(dim=1)]
        nb = torch.unique(torch.cat([nb, cand]), dim=0)
        if ring < k_ring - 1:
            front = nb
    lin = (nb[:, 0] * R + nb[:, 1]) * R + nb[:, 2]

    def member(s):
        if len(s) == 0:
            return torch.zeros(len(lin), dtype=torch.bool, device=device)
        sl, _ = torch.sort((s[:, 0].long() * R + s[:, 1].long()) * R + s[:, 2].long())
        i = torch.searchsorted(sl, lin).clamp(max=len(sl) - 1)
        return sl[i] == lin

    signs = torch.full((len(lin),), SIGN_WATER, dtype=torch.int8, device=device)
    signs[member(dry_ijk)] = SIGN_DRY
    signs[member(water_ijk)] = SIGN_WATER
    signs[member(dam_ijk)] = 0
    return nb.int(), signs


def sparse_flood_fill(bvh: MeshBVH, resolution: int, source: Tuple[int, int, int] = (0, 0, 0),
                      connectivity: int = 26, min_level: int = 0, k_ring: int = 1,
                      bounds: Optional[torch.Tensor] = None, verbose: bool = True) -> Dict[str, torch.Tensor]:
    """Inside/outside classification of the voxels of a mesh at ``resolution``^3.

    Returns dam_ijk, water_ijk, dry_ijk, tide_ijk (collision voxels),
    extended_tide_ijk and dry_adjacent_ijk (the dam's k-ring neighbours signed
    water / dry)."""
    if k_ring > _PADDING:
        raise NotImplementedError(f"k_ring > {_PADDING} reaches outside the flood crop")
    device = bvh.device
    max_level = int(np.log2(resolution))
    octree = OctreeIndexer(max_level=max_level, bounds=bounds, device=device)
    vox = octree.octree_traverse(bvh, min_level=min_level)
    t0 = time.time()
    out = sparse_flood_fill_from_voxels(vox, resolution, source, connectivity, k_ring)
    if verbose:
        logger.info(f"Sparse Flood Fill completed in {time.time() - t0:.4f}s")
    return out


def sparse_flood_fill_from_voxels(vox: torch.Tensor, resolution: int, source: Tuple[int, int, int] = (0, 0, 0),
                                  connectivity: int = 26, k_ring: int = 1) -> Dict[str, torch.Tensor]:
    """sparse_flood_fill's output dict for given surface voxels ([N, 3] ijk
    on the device; the octree's level-``log2(resolution)`` voxels):
    the flood, classification and k-ring steps of
    apps/sparse_flood_fill.py:134-166 without the octree."""
    if k_ring > _PADDING:
        raise NotImplementedError(f"k_ring > {_PADDING} reaches outside the flood crop")
    e3 = torch.empty(0, 3, dtype=torch.int32, device=vox.device)
    st = flood_planes(vox, resolution, source, _PADDING, connectivity, k_ring) if len(vox) else None
    if st is None:
        return {k: e3.clone() for k in ("dam_ijk", "water_ijk", "dry_ijk", "tide_ijk", "extended_tide_ijk",
                                         "dry_adjacent_ijk")}
    dt = torch.promote_types(torch.int32, vox.dtype)
    dam = _out(st.coords("dam"), dt)
    return {
        "dam_ijk": dam,
        "water_ijk": _out(st.coords("water"), dt),
        "dry_ijk": _out(st.coords("dry"), dt),
        "tide_ijk": _out(st.coords("collision"), dt),
        "extended_tide_ijk": st.coords("ext_tide").int() if len(dam) and k_ring > 0 else e3.clone(),
        "dry_adjacent_ijk": st.coords("dry_adj").int() if len(dam) and k_ring > 0 else e3.clone(),
    }


def _check_cells(t: torch.Tensor, resolution: int, name: str) -> None:
    if len(t):
        lo, hi = (int(x) for x in torch.aminmax(t.reshape(-1)))
        if lo < 0 or hi >= resolution:
            raise IndexError(f"{name}: coordinates [{lo}, {hi}] outside the {resolution}^3 grid")


def _wet(kind: int, dam_ijk: torch.Tensor, collision_ijk: torch.Tensor, resolution: int) -> torch.Tensor:
    per = (6, 12, 8)[kind]
    if len(dam_ijk) == 0:
        return torch.zeros((0, per), dtype=torch.bool, device=dam_ijk.device)
    _check_cells(collision_ijk, resolution, "collision_ijk")
    return wet_mask(kind, dam_ijk, collision_ijk, resolution)


def get_dam_wet_faces(dam_ijk: torch.Tensor, collision_ijk: torch.Tensor, resolution: int) -> torch.Tensor:
    """bool [N,6]: dam voxel's 6-neighbour (NEIGHBOR_OFFSETS_6 order, in
    range) is a collision voxel (sparse_flood_fill.py:169-185)."""
    return _wet(0, dam_ijk, collision_ijk, resolution)


def get_dam_wet_edges(dam_ijk: torch.Tensor, collision_ijk: torch.Tensor, resolution: int) -> torch.Tensor:
    """bool [N,12]: any of an edge's EDGE_NEIGHBOR_OFFSETS neighbours is a
    collision voxel (sparse_flood_fill.py:188-206)."""
    return _wet(1, dam_ijk, collision_ijk, resolution)


def get_dam_wet_corners(dam_ijk: torch.Tensor, collision_ijk: torch.Tensor, resolution: int) -> torch.Tensor:
    """bool [N,8]: any of a corner's CORNER_NEIGHBOR_OFFSETS neighbours is a
    collision voxel (sparse_flood_fill.py:209-227)."""
    return _wet(2, dam_ijk, collision_ijk, resolution)


def construct_surface_dual_graph(dam_ijk: torch.Tensor, collision_ijk: torch.Tensor,
                                 dry_adjacent_ijk: Optional[torch.Tensor], octree: OctreeIndexer):
    """Surface dual graph of the dam, collision and dry-adjacent voxels
    (sparse_flood_fill.py:230-291).

    Returns (voxelgrid_vertices [V,3] float32 world positions, vertices_ijk
    [V,3] int64, cube_idx [N,8] int64, adj_idx [N,6] int64 (-1 = none),
    active_cells_ijk [N,3]): cells and vertices in lexicographic order."""
    device = dam_ijk.device if len(dam_ijk) > 0 else (collision_ijk.device if len(collision_ijk) > 0
                                                      else torch.device("cuda"))
    if dry_adjacent_ijk is None:
        dry_adjacent_ijk = torch.empty((0, 3), dtype=torch.int32, device=device)
    R = 2 ** octree.max_level
    parts = [c for c in (dam_ijk, collision_ijk, dry_adjacent_ijk) if len(c) > 0]
    if not parts:
        raise RuntimeError("construct_surface_dual_graph: dam, collision and dry-adjacent voxel lists are all empty")
    cells = torch.cat(parts, dim=0)
    active = unique_cells(cells, R).to(cells.dtype)
    cell = ((octree.bounds[1] - octree.bounds[0]) / float(R)).tolist()
    cube_idx, _, vijk, world = cell_vertices(active, R, lo=octree.bounds[0].tolist(), cell=cell, want_ijk=True)
    adj = cell_neighbors(active, R)
    return world, vijk, cube_idx, adj, active


__all__ = ["sparse_flood_fill", "get_dam_neighborhood_signs", "get_dam_wet_faces", "get_dam_wet_edges",
           "get_dam_wet_corners", "construct_surface_dual_graph"]

"""Oracle (test infrastructure only): Voxelizer.voxelize_surface /
voxelize_with_faces restated with numpy and the C++ SAT oracle, as the
checker of atom3d_amd.apps.voxelizer.

Follows Atom3D:
  candidate block (floor/ceil of grid coords, clamp, min/max)  grid/cube_grid.py:542-557
  cell boxes float(ijk) * cell + b0, + cell                    grid/cube_grid.py:164-203
  SAT hit mask and (box, face) pairs                           kernels/bvh_kernels.cu:285-353, 555-610
  CSR from the pairs (unique, bincount, cumsum)                apps/voxelizer.py:75-131
"""
from __future__ import annotations

import numpy as np

from . import SahBvh


def candidate_block(V, F, b0, cell, res):
    """(origin [3], dims [3]) of the cells covering the faces' boxes."""
    t = np.asarray(V, np.float32)[np.asarray(F, np.int64)]
    b0, cell = np.asarray(b0, np.float32), np.asarray(cell, np.float32)
    gmin = np.clip(np.floor((t.min(1) - b0) / cell).astype(np.int64), 0, res - 1)
    gmax = np.clip(np.ceil((t.max(1) - b0) / cell).astype(np.int64), 0, res - 1)
    lo, hi = gmin.min(0), gmax.max(0)
    return lo, hi - lo + 1


def block_cells(origin, dims):
    ax = [np.arange(o, o + d, dtype=np.int64) for o, d in zip(origin, dims)]
    return np.stack(np.meshgrid(*ax, indexing="ij"), -1).reshape(-1, 3)


def cell_boxes(ijk, b0, cell):
    lo = (np.asarray(ijk).astype(np.float32) * np.asarray(cell, np.float32)) + np.asarray(b0, np.float32)
    return lo, lo + np.asarray(cell, np.float32)


def voxelize_with_faces(V, F, b0, cell, res, strategy="candidate"):
    """(voxel ijk [K,3] int64, face_indices, face_start, face_count) with the
    faces of a voxel ascending; strategy "all" tests every cell."""
    if strategy == "all":
        origin, dims = np.zeros(3, np.int64), np.full(3, res, np.int64)
    else:
        origin, dims = candidate_block(V, F, b0, cell, res)
    cells = block_cells(origin, dims)
    lo, hi = cell_boxes(cells, b0, cell)
    _, pa, pf = SahBvh(V, F).aabb(lo, hi)
    order = np.lexsort((pf, pa))  # (cell, face) order
    pa, pf = pa[order], pf[order]
    uniq, inv = np.unique(pa, return_inverse=True)
    count = np.bincount(inv, minlength=len(uniq)).astype(np.int32)
    start = (np.cumsum(count) - count).astype(np.int32)
    return cells[uniq], pf.astype(np.int32), start, count

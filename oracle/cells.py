"""Oracle (test infrastructure only): the voxel-to-dual-graph glue restated
with numpy, as the checker of atom3d_amd.kernels.cells.

Follows Atom3D:
  get_dam_wet_faces / _edges / _corners   apps/sparse_flood_fill.py:169-227
  construct_surface_dual_graph            apps/sparse_flood_fill.py:230-291
  CubeGrid.voxel_unique_vertices          grid/cube_grid.py:237-260
Neighbour tables come from atom3d_amd.constants, whose values
tests/test_oracle_cells.py pins to the reference's (digest of its literal
tables).
"""
from __future__ import annotations

import numpy as np

from atom3d_amd.constants import CORNER_NEIGHBOR_OFFSETS, CUBE_CORNERS, EDGE_NEIGHBOR_OFFSETS, NEIGHBOR_OFFSETS_6

_CORNERS = CUBE_CORNERS.numpy().astype(np.int64)


def unique_cells(ijk: np.ndarray) -> np.ndarray:
    """Lexicographically sorted unique rows (torch.unique(dim=0))."""
    return np.unique(np.asarray(ijk).reshape(-1, 3), axis=0)


def cell_vertices(ijk: np.ndarray):
    """(vertices_ijk [V,3] int64 sorted, cube_vert [N,8] int64): the unique
    corners of the cells and each cell corner's row among them."""
    c = np.asarray(ijk, np.int64).reshape(-1, 3)
    cand = (c[:, None, :] + _CORNERS[None]).reshape(-1, 3)
    v, inv = np.unique(cand, axis=0, return_inverse=True)
    return v, inv.reshape(-1, 8).astype(np.int64)


def world(vijk: np.ndarray, lo, cell) -> np.ndarray:
    """lo + ijk.float() * cell in float32 (two roundings)."""
    return (np.asarray(vijk).astype(np.float32) * np.asarray(cell, np.float32)) + np.asarray(lo, np.float32)


def cell_neighbors(cells: np.ndarray, R: int) -> np.ndarray:
    """[N,6] index of the cell whose flat key (i*R + j)*R + k equals the
    neighbour's unchecked flat key, else -1 (sparse_flood_fill.py:273-283)."""
    c = np.asarray(cells, np.int64)
    keys = (c[:, 0] * R + c[:, 1]) * R + c[:, 2]
    order = np.argsort(keys, kind="stable")
    sk = keys[order]
    nb = c[:, None, :] + NEIGHBOR_OFFSETS_6.numpy().astype(np.int64)[None]
    want = (nb[..., 0] * R + nb[..., 1]) * R + nb[..., 2]
    pos = np.clip(np.searchsorted(sk, want), 0, max(len(sk) - 1, 0))
    found = sk[pos] == want if len(sk) else np.zeros(want.shape, bool)
    return np.where(found, order[pos], -1).astype(np.int64)


def construct_surface_dual_graph(dam, collision, dry_adjacent, R: int, lo, hi):
    """(world [V,3] f32, vertices_ijk [V,3], cube_idx [N,8], adj [N,6],
    active cells [N,3]) for bounds lo..hi."""
    parts = [np.asarray(p).reshape(-1, 3) for p in (dam, collision, dry_adjacent) if p is not None and len(p)]
    cells = unique_cells(np.concatenate(parts))
    v, cube = cell_vertices(cells)
    lo32 = np.asarray(lo, np.float32)
    cell = (np.asarray(hi, np.float32) - lo32) / np.float32(R)
    return world(v, lo32, cell), v, cube, cell_neighbors(cells, R), cells


def _wet(dam, collision, R, offsets):
    """offsets [K, M, 3]: out[n, k] = any m: dam[n] + offsets[k, m] in range
    and in the collision set."""
    mask = np.zeros((R, R, R), bool)
    c = np.asarray(collision, np.int64).reshape(-1, 3)
    if len(c):
        mask[c[:, 0], c[:, 1], c[:, 2]] = True
    d = np.asarray(dam, np.int64).reshape(-1, 3)
    nb = d[:, None, None, :] + offsets[None]
    ok = ((nb >= 0) & (nb < R)).all(-1)
    nbc = np.where(ok[..., None], nb, 0)
    hit = mask[nbc[..., 0], nbc[..., 1], nbc[..., 2]] & ok
    return hit.any(-1)


def wet_faces(dam, collision, R):
    return _wet(dam, collision, R, NEIGHBOR_OFFSETS_6.numpy().astype(np.int64)[:, None, :])


def wet_edges(dam, collision, R):
    return _wet(dam, collision, R, EDGE_NEIGHBOR_OFFSETS.numpy().astype(np.int64))


def wet_corners(dam, collision, R):
    return _wet(dam, collision, R, CORNER_NEIGHBOR_OFFSETS.numpy().astype(np.int64))

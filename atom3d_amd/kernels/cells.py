"""Voxel-to-dual-graph glue on the GPU (C ABI a3d_unique_cells,
a3d_cell_vertices(_get), a3d_cell_neighbors, a3d_wet_mask).

The integer formats between the flood fill / octree and the SDF / DMC steps:
unique cells, their unique corner vertices with the cube -> vertex table,
the 6-neighbour table and the wet face / edge / corner masks of
Atom3D atom3d/apps/sparse_flood_fill.py:169-291 and
atom3d/grid/cube_grid.py:237-260.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence, Tuple

import torch

from . import _lib


def _ijk32(t: torch.Tensor, name: str) -> torch.Tensor:
    _lib.require_cuda(t, name)
    if t.dim() != 2 or t.shape[1] != 3:
        raise ValueError(f"{name} must be [N, 3], got {tuple(t.shape)}")
    if t.dtype.is_floating_point or t.dtype == torch.bool:
        raise TypeError(f"{name} must hold integers, got {t.dtype}")
    return t.to(torch.int32).contiguous()


# method: 0 = the library chooses (bitmaps up to 2048^3 when the grid is not
# too sparse, else a radix sort), 1 = sort, 2 = bitmaps; results are the same.
def _ws(n: int, resolution: int, method: int, device) -> Tuple[torch.Tensor, int]:
    nb = ctypes.c_int64()
    _lib.call("a3d_cells_ws_bytes", int(n), int(resolution), int(method), ctypes.byref(nb))
    return torch.empty(max(nb.value, 1), dtype=torch.uint8, device=device), nb.value


def unique_cells(ijk: torch.Tensor, resolution: int, method: int = 0) -> torch.Tensor:
    """Sorted (lexicographic) unique rows of int cells [N,3] -> int32 [U,3]."""
    c = _ijk32(ijk, "cells")
    dev = c.device
    n = int(c.shape[0])
    out = torch.empty((max(n, 1), 3), dtype=torch.int32, device=dev)
    if n == 0:
        return out[:0]
    ws, nb = _ws(n, resolution, method, dev)
    cnt = ctypes.c_int64()
    _lib.call("a3d_unique_cells", _lib.ptr(c), n, int(resolution), int(method), _lib.ptr(ws), nb, _lib.ptr(out),
              ctypes.byref(cnt), _lib.stream(dev))
    return out[:cnt.value]


def cell_vertices(ijk: torch.Tensor, resolution: int, lo: Optional[Sequence[float]] = None,
                  cell: Optional[Sequence[float]] = None, want_keys: bool = False, want_ijk: bool = False,
                  method: int = 0):
    """Unique corner vertices of cells [N,3] (duplicates allowed).

    Returns (cube_vert int64 [N,8], vkey int64 [V] | None, vijk int64 [V,3] |
    None, world float32 [V,3] | None): vertices in lexicographic order,
    ``cube_vert[n, k]`` the rank of corner k (CUBE_CORNERS order) of cell n,
    ``vkey`` flat indices in the (resolution+1)^3 vertex grid and ``world`` =
    ijk * cell + lo when ``lo`` and ``cell`` are given."""
    c = _ijk32(ijk, "cells")
    dev = c.device
    n = int(c.shape[0])
    cube_vert = torch.empty((n, 8), dtype=torch.int64, device=dev)
    nv = ctypes.c_int64()
    ws, nb = _ws(max(n, 1), resolution, method, dev)
    if n:
        _lib.call("a3d_cell_vertices", _lib.ptr(c), n, int(resolution), int(method), _lib.ptr(ws), nb,
                  _lib.ptr(cube_vert), ctypes.byref(nv), _lib.stream(dev))
    V = nv.value
    vkey = torch.empty((V,), dtype=torch.int64, device=dev) if want_keys else None
    vijk = torch.empty((V, 3), dtype=torch.int64, device=dev) if want_ijk else None
    world = torch.empty((V, 3), dtype=torch.float32, device=dev) if lo is not None else None
    if V and (vkey is not None or vijk is not None or world is not None):
        lo_c = (ctypes.c_float * 3)(*[float(x) for x in lo]) if world is not None else None
        ce_c = (ctypes.c_float * 3)(*[float(x) for x in cell]) if world is not None else None
        _lib.call("a3d_cell_vertices_get", _lib.ptr(ws), n, int(resolution), int(method), V, _lib.ptr(vkey),
                  _lib.ptr(vijk), _lib.ptr(world), lo_c, ce_c, _lib.stream(dev))
    return cube_vert, vkey, vijk, world


def cell_neighbors(cells_sorted: torch.Tensor, resolution: int, method: int = 0) -> torch.Tensor:
    """int64 [N,6]: index of each sorted unique cell's neighbour at +x,-x,+y,
    -y,+z,-z (unchecked flat key, as the reference looks it up), else -1."""
    c = _ijk32(cells_sorted, "cells")
    n = int(c.shape[0])
    adj = torch.empty((n, 6), dtype=torch.int64, device=c.device)
    if n:
        ws, nb = _ws(n, resolution, method, c.device)
        _lib.call("a3d_cell_neighbors", _lib.ptr(c), n, int(resolution), int(method), _lib.ptr(ws), nb,
                  _lib.ptr(adj), _lib.stream(c.device))
    return adj


_WET_PER = {0: 6, 1: 12, 2: 8}


def wet_mask(kind: int, dam_ijk: torch.Tensor, collision_ijk: torch.Tensor, resolution: int) -> torch.Tensor:
    """bool [N, 6 | 12 | 8] (kind 0 faces, 1 edges, 2 corners): whether an
    in-range neighbour of each dam voxel is a collision voxel."""
    d = _ijk32(dam_ijk, "dam_ijk")
    dev = d.device
    n = int(d.shape[0])
    out = torch.empty((n, _WET_PER[kind]), dtype=torch.uint8, device=dev)
    if n == 0:
        return out.bool()
    cc = _ijk32(collision_ijk.to(dev), "collision_ijk") if len(collision_ijk) else None
    nb = ctypes.c_int64()
    _lib.call("a3d_wet_ws_bytes", int(resolution), ctypes.byref(nb))
    bitmap = torch.empty(nb.value, dtype=torch.uint8, device=dev)
    _lib.call("a3d_wet_mask", int(kind), _lib.ptr(d), n, _lib.ptr(cc), 0 if cc is None else int(cc.shape[0]),
              int(resolution), _lib.ptr(bitmap), nb.value, _lib.ptr(out), _lib.stream(dev))
    return out.bool()


__all__ = ["unique_cells", "cell_vertices", "cell_neighbors", "wet_mask"]

"""Oracle (test infrastructure only): the reference's sparse flood-fill
pipeline restated with numpy around the C++ dense BFS.

Follows Atom3D atom3d/kernels/flood_fill.py:118-258 (crop = dam bounding box
+- padding clamped to the grid, occupied-start guard, extended dam, nonzero
extraction) and atom3d/apps/sparse_flood_fill.py:34-166 (k-ring
neighbourhood signs: dam 0, water 1, dry -1, default water).
"""
from __future__ import annotations

import numpy as np

from . import lib, _SIGS  # noqa: F401
import ctypes

_u8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_i = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def _bind():
    L = lib()
    f = L.orc_flood_fill_dense
    f.restype = ctypes.c_int64
    f.argtypes = [_u8, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                  ctypes.c_int, _i]
    return f


def flood_fill_dense(occ: np.ndarray, start, conn: int = 26) -> np.ndarray:
    """Dense labels int32[D,H,W]: -2 dry, -1 dam, 1 collision, 2 water."""
    occ = np.ascontiguousarray(occ, np.uint8)
    D, H, W = occ.shape
    mask = np.empty(occ.shape, np.int32)
    _bind()(occ.reshape(-1), D, H, W, int(start[0]), int(start[1]), int(start[2]), conn, mask.reshape(-1))
    return mask


def _dilate(x: np.ndarray, conn: int) -> np.ndarray:
    """max over the 3x3x3 cube (18, 26) or the 6-cross plus centre (6)."""
    p = np.pad(x, 1)
    D, H, W = x.shape
    out = np.zeros_like(x)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                if conn == 6 and abs(dz) + abs(dy) + abs(dx) > 1:
                    continue
                out |= p[1 + dz:1 + dz + D, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W]
    return out


def flood_fill_sparse(dam_coords: np.ndarray, resolution: int, source=(0, 0, 0), padding: int = 2,
                      connectivity: int = 26):
    """(water, dry, collision, true_dam, dam_mask, extended_dam), coordinates
    as int64 [n,3] in lexicographic order."""
    dam = np.asarray(dam_coords, np.int64).reshape(-1, 3)
    e3 = np.zeros((0, 3), np.int64)
    if len(dam) == 0:
        return e3, e3, e3, e3, np.zeros(0, bool), e3
    lo = np.maximum(dam.min(0) - padding, 0)
    hi = np.minimum(dam.max(0) + padding, resolution - 1)
    size = hi - lo + 1
    occ = np.zeros(size, bool)
    loc = dam - lo
    occ[loc[:, 0], loc[:, 1], loc[:, 2]] = True
    ls = np.asarray(source, np.int64) - lo
    if np.all(ls >= 0) and np.all(ls < size):
        if occ[tuple(ls)]:
            return e3, e3, e3, e3, np.zeros(0, bool), e3
        start = ls
    else:
        start = np.zeros(3, np.int64)
        if occ[0, 0, 0]:
            return e3, e3, e3, e3, np.zeros(0, bool), e3
    mask = flood_fill_dense(occ, start, connectivity)
    dil = _dilate(mask > 0, 6 if connectivity == 6 else 26)
    ext = (mask == -2) & occ & dil
    mask[ext] = -1

    def nz(m):
        return np.argwhere(m).astype(np.int64) + lo

    dam_mask = mask[loc[:, 0], loc[:, 1], loc[:, 2]] == -1
    return nz(mask == 2), nz(mask == -2), nz(mask == 1), nz(mask == -1), dam_mask, nz(ext)


def neighborhood_signs(dam, water, dry, resolution: int, k_ring: int = 1):
    """k-ring 26-neighbourhood of the dam (unique, lexicographic) and signs:
    dam 0, else water 1, else dry -1, else 1."""
    dam = np.asarray(dam, np.int64).reshape(-1, 3)
    if len(dam) == 0:
        return np.zeros((0, 3), np.int64), np.zeros(0, np.int8)
    offs = np.array([[dx, dy, dz] for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)
                     if (dx, dy, dz) != (0, 0, 0)], np.int64)
    allnb = dam
    front = dam
    for ring in range(k_ring):
        nb = (front[:, None, :] + offs[None]).reshape(-1, 3)
        nb = nb[np.all((nb >= 0) & (nb < resolution), axis=1)]
        allnb = np.unique(np.concatenate([allnb, nb]), axis=0)
        if ring < k_ring - 1:
            front = allnb
    R = resolution
    lin = allnb[:, 0] * R * R + allnb[:, 1] * R + allnb[:, 2]

    def member(s):
        s = np.asarray(s, np.int64).reshape(-1, 3)
        if len(s) == 0:
            return np.zeros(len(lin), bool)
        return np.isin(lin, s[:, 0] * R * R + s[:, 1] * R + s[:, 2])

    signs = np.ones(len(allnb), np.int8)
    signs[member(dry)] = -1
    signs[member(water)] = 1
    signs[member(dam)] = 0
    return allnb, signs


def sparse_flood_fill_from_voxels(voxels, resolution: int, source=(0, 0, 0), connectivity: int = 26,
                                  k_ring: int = 1):
    """The output dict of sparse_flood_fill given the octree voxels."""
    water, dry, col, dam, _, _ = flood_fill_sparse(voxels, resolution, source, padding=10,
                                                   connectivity=connectivity)
    nb, s = neighborhood_signs(dam, water, dry, resolution, k_ring)
    return {"dam_ijk": dam, "water_ijk": water, "dry_ijk": dry, "tide_ijk": col,
            "extended_tide_ijk": nb[s == 1], "dry_adjacent_ijk": nb[s == -1]}

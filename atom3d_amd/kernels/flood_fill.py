"""Flood fill over a voxel crop (HIP bit-plane fixpoint).

Same functions, arguments and results as Atom3D atom3d/kernels/flood_fill.py:
``flood_fill_3d`` (:76-115) and ``flood_fill_3d_sparse`` (:118-258).  The
labels come from the library's bit-plane fixpoint instead of a host-driven
BFS; they equal the reference's (they depend on reachability only).
"""
from __future__ import annotations

import ctypes
from typing import Tuple

import torch

from . import _lib

PLANE = {"occ": 0, "reached": 1, "water": 2, "collision": 3, "dam": 4, "dry": 5, "extended": 6,
         "neighborhood": 7, "ext_tide": 8, "dry_adj": 9}
NPLANES = 11


def floodfill_available() -> bool:
    try:
        _lib.lib()
    except Exception:
        return False
    return torch.cuda.is_available()


class FloodState:
    """Bit planes of one flood-filled crop (see include/atom3d_amd.h)."""

    def __init__(self, dims, lo, device):
        self.D, self.H, self.W = (int(x) for x in dims)
        self.lo = [int(x) for x in lo]
        self.device = device
        pw, wsb = ctypes.c_int64(), ctypes.c_int64()
        _lib.call("a3d_flood_sizes", self.D, self.H, self.W, ctypes.byref(pw), ctypes.byref(wsb))
        self.plane_words = pw.value
        self.state = torch.empty(NPLANES * pw.value, dtype=torch.int32, device=device)
        self.ws = torch.empty(wsb.value, dtype=torch.uint8, device=device)
        self.ws_bytes = wsb.value
        self.totals = [0] * NPLANES
        self.passes = 0

    def _lo(self):
        return (ctypes.c_int64 * 3)(*self.lo)

    def run(self, dam_ijk, start, conn: int, k_ring: int) -> bool:
        """False when the start voxel is occupied."""
        tot = (ctypes.c_int64 * NPLANES)()
        passes = ctypes.c_int64()
        N = 0 if dam_ijk is None else int(dam_ijk.shape[0])
        with torch.cuda.device(self.device):
            _lib.call("a3d_flood_fill", _lib.ptr(dam_ijk), N, self._lo(), self.D, self.H, self.W,
                      int(start[0]), int(start[1]), int(start[2]), int(conn), int(k_ring), _lib.ptr(self.state),
                      _lib.ptr(self.ws), self.ws_bytes, tot, ctypes.byref(passes), _lib.stream(self.device))
        self.totals = list(tot)
        self.passes = passes.value
        return self.passes >= 0

    def coords(self, name: str) -> torch.Tensor:
        p = PLANE[name]
        n = self.totals[p]
        out = torch.empty(n, 3, dtype=torch.int64, device=self.device)
        if n:
            with torch.cuda.device(self.device):
                _lib.call("a3d_flood_extract", _lib.ptr(self.state), self.D, self.H, self.W, p, self._lo(),
                          _lib.ptr(out), _lib.ptr(self.ws), self.ws_bytes, _lib.stream(self.device))
        return out

    def dam_mask(self, dam_ijk: torch.Tensor) -> torch.Tensor:
        out = torch.empty(dam_ijk.shape[0], dtype=torch.uint8, device=self.device)
        with torch.cuda.device(self.device):
            _lib.call("a3d_flood_dam_mask", _lib.ptr(self.state), self.D, self.H, self.W, _lib.ptr(dam_ijk),
                      dam_ijk.shape[0], self._lo(), _lib.ptr(out), _lib.stream(self.device))
        return out.view(torch.bool)


def flood_fill_3d(occupancy: torch.Tensor, start_point: Tuple[int, int, int] = (0, 0, 0),
                  connectivity: int = 26) -> torch.Tensor:
    """Dense labels int32[D,H,W] (-2 dry, -1 dam, 1 collision, 2 water) of a
    bool occupancy grid, flooding from start_point = (d, h, w) indices."""
    if not occupancy.is_cuda:
        raise ValueError("occupancy must be on CUDA device")
    if occupancy.dim() != 3:
        raise ValueError(f"Expected 3D tensor, got {occupancy.dim()}D")
    occ = occupancy.contiguous().to(torch.bool)
    D, H, W = occ.shape
    dev = occ.device
    mask = torch.empty(D, H, W, dtype=torch.int32, device=dev)
    st = FloodState((D, H, W), (0, 0, 0), dev)
    z, y, x = (int(v) for v in start_point)
    if not (0 <= x < W and 0 <= y < H and 0 <= z < D):
        return mask.fill_(-2)  # the reference never seeds an out-of-range start
    with torch.cuda.device(dev):
        _lib.call("a3d_flood_pack_dense", _lib.ptr(occ), D, H, W, _lib.ptr(st.state), _lib.stream(dev))
    if not st.run(None, (z, y, x), connectivity, 0):
        # occupied start: the reference marks it water and floods from it
        raise ValueError(f"start point {tuple(start_point)} is occupied")
    with torch.cuda.device(dev):
        _lib.call("a3d_flood_dense_mask", _lib.ptr(st.state), D, H, W, _lib.ptr(mask), _lib.stream(dev))
    return mask


def _empty_result(device):
    e = torch.empty(0, 3, dtype=torch.int32, device=device)
    return e, e.clone(), e.clone(), e.clone(), torch.empty(0, dtype=torch.bool, device=device), e.clone()


def _crop(dam_coords: torch.Tensor, resolution: int, source, padding: int):
    b = torch.empty(6, dtype=torch.int64, device=dam_coords.device)
    with torch.cuda.device(dam_coords.device):
        _lib.call("a3d_ijk_bounds", _lib.ptr(dam_coords), int(dam_coords.shape[0]), _lib.ptr(b),
                  _lib.stream(dam_coords.device))
    b = b.tolist()
    lo_l = [max(v - padding, 0) for v in b[:3]]
    hi_l = [min(v + padding, resolution - 1) for v in b[3:]]
    size = [h - l + 1 for l, h in zip(lo_l, hi_l)]
    ls = [int(s) - l for s, l in zip(source, lo_l)]
    inside = all(0 <= a < n for a, n in zip(ls, size))
    return lo_l, size, (ls if inside else [0, 0, 0]), inside


def flood_planes(dam_coords: torch.Tensor, resolution: int, source, padding: int, connectivity: int, k_ring: int):
    """Run the flood on the crop around dam_coords; None when the start voxel
    is occupied (the reference's empty result)."""
    dam = dam_coords.to(torch.int64).contiguous()
    lo, size, start, inside = _crop(dam, resolution, source, padding)
    st = FloodState(size, lo, dam.device)
    if not st.run(dam, start, connectivity, k_ring):
        where = f"Start point {tuple(source)}" if inside else "Default start (0,0,0)"
        print(f"[Flood Fill Warning] {where} is OCCUPIED. Returning empty result.")
        return None
    return st


def _out(t: torch.Tensor, dtype) -> torch.Tensor:
    return t.to(dtype) if t.shape[0] > 0 else torch.empty(0, 3, dtype=torch.int32, device=t.device)


def flood_fill_3d_sparse(dam_coords: torch.Tensor, resolution: int, source: Tuple[int, int, int] = (0, 0, 0),
                         padding: int = 2, connectivity: int = 26, return_water_dry: bool = True):
    """(water, dry, collision, true_dam, dam_mask, extended_dam) of the flood
    on the crop around dam_coords; coordinates in lexicographic order, in the
    dtype the reference produces (int32 + dam_coords' dtype; int32 if empty)."""
    device = dam_coords.device
    if len(dam_coords) == 0:
        return _empty_result(device)
    st = flood_planes(dam_coords, resolution, source, padding, connectivity, 0)
    if st is None:
        return _empty_result(device)
    dt = torch.promote_types(torch.int32, dam_coords.dtype)
    e = torch.empty(0, 3, dtype=torch.int32, device=device)
    water = _out(st.coords("water"), dt) if return_water_dry else e
    dry = _out(st.coords("dry"), dt) if return_water_dry else e.clone()
    dam = dam_coords.to(torch.int64).contiguous()
    return (water, dry, _out(st.coords("collision"), dt), _out(st.coords("dam"), dt), st.dam_mask(dam),
            _out(st.coords("extended"), dt))


__all__ = ["floodfill_available", "flood_fill_3d", "flood_fill_3d_sparse", "flood_planes", "FloodState"]

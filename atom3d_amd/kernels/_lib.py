"""ctypes binding of the gfx950 HIP library ``libatom3d_amd.so``.

Replaces the reference's JIT-built pybind11 modules (Atom3D
atom3d/kernels/bvh.py:21-51, kernels/__init__.py:15-83,
kernels/flood_fill.py:20-62).  The library is built ahead of time in-tree
(``python -m atom3d_amd.build``); there is no JIT step and no fallback: if the
library is missing or a call fails, a RuntimeError says why.

Calls go through ``ctypes.CDLL`` (the GIL is released during each call) and
pass raw device pointers plus the caller's current HIP stream, so every kernel
is ordered with the surrounding PyTorch work.
"""
from __future__ import annotations

import ctypes
import os
import threading

import torch

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "_lib", "libatom3d_amd.so")

_lib = None
_lock = threading.Lock()

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int
_f32 = ctypes.c_float
_f64 = ctypes.c_double
_u64 = ctypes.c_uint64
_pf32 = ctypes.POINTER(ctypes.c_float)
_pi64 = ctypes.POINTER(ctypes.c_int64)

# name -> argtypes.  Every entry point returns int (0 = OK).
SIGNATURES: dict[str, list] = {
    "a3d_bvh_sizes": [_i64, _i32, _pi64, _pi64],
    "a3d_bvh_build": [_vp, _i64, _vp, _i64, _i32, _vp, _i64, _vp, _i64, _vp],
    "a3d_ray_nearest": [_vp, _vp, _vp, _i64, _f32, _vp, _vp, _vp, _vp, _vp, _vp],
    "a3d_ray_all_count": [_vp, _vp, _vp, _i64, _f32, _vp, _vp],
    "a3d_ray_all_hits": [_vp, _vp, _vp, _i64, _f32, _vp, _vp, _vp, _vp, _vp, _vp],
    "a3d_ray_eval": [_vp, _i64, _vp, _i64, _vp, _vp, _i64, _vp, _vp, _f32, _vp, _vp, _vp, _vp, _vp],
    "a3d_ray_crossings": [_vp, _vp, _vp, _i64, _f32, _vp, _vp],
    "a3d_ray_crossings_dirs": [_vp, _vp, _i64, _vp, _i32, _f32, _f32, _vp, _vp],
    "a3d_bvh_build_ex": [_vp, _i64, _vp, _i64, _i32, _i32, _vp, _i64, _vp, _i64, _vp],
    "a3d_bvh_info": [_vp, _pi64, _vp],
    "a3d_sah_sizes": [_i64, _i32, _pi64, _pi64],
    "a3d_sah_build": [_vp, _i64, _vp, _i64, _i32, _vp, _i64, _vp, _i64, ctypes.POINTER(ctypes.c_int32), _vp],
    "a3d_query_order_ws_bytes": [_i64, _pi64],
    "a3d_query_order": [_vp, _i64, _vp, _vp, _i64, _vp],
    "a3d_query_order_noalloc": [_vp, _i64, _vp, _vp],
    "a3d_query_order_work_ws_bytes": [_i64, _pi64],
    "a3d_query_order_work": [_vp, _i64, _vp, _i32, _pf32, _f32, _vp, _vp, _i64, _vp],
    "a3d_bvh_udf": [_vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _i32, _vp, _vp],
    "a3d_bvh_udf_work": [_vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _i32, _vp, _vp],
    "a3d_bvh_udf_seeded": [_vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _vp],
    "a3d_grid_points": [ctypes.POINTER(ctypes.c_float), _f32, _i32, _vp, _vp, _vp],
    "a3d_udf_grid_seed": [_vp, _vp, _vp, _vp, _i32, _vp, _f32, _vp, _vp, _i64, _f32, _vp, _vp],
    "a3d_bvh_udf_grid": [_vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _i32, _vp, _f32, _f32, _vp, _vp, _vp, _vp, _i32,
                         _vp],
    "a3d_point_dist": [_vp, _vp, _i64, _vp, _vp],
    "a3d_point_dist_backward": [_vp, _vp, _vp, _vp, _i64, _vp, _vp],
    "a3d_sdf_sign": [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "a3d_bvh_aabb_hit": [_vp, _vp, _vp, _i64, _vp, _vp],
    "a3d_bvh_aabb_count": [_vp, _vp, _vp, _i64, _vp, _vp],
    "a3d_bvh_aabb_pairs": [_vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp],
    "a3d_pairs_sort_ws_bytes": [_i64, _pi64],
    "a3d_pairs_sort": [_vp, _vp, _i64, _vp, _i64, _vp],
    "a3d_sat_clip": [_vp, _vp, _vp, _vp, _vp, _i64, _i32, _f32, _vp, _vp, _vp, _vp, _vp, _vp],
    "a3d_octree_ws_bytes": [_i64, _pi64],
    "a3d_octree_level": [_vp, ctypes.POINTER(ctypes.c_float), _i32, _vp, _i64, _vp, _vp, _vp, _i64, _vp],
    "a3d_octree_unpack": [_vp, _i64, _vp, _vp],
    "a3d_flood_sizes": [_i32, _i32, _i32, _pi64, _pi64],
    "a3d_flood_pack_dense": [_vp, _i32, _i32, _i32, _vp, _vp],
    "a3d_flood_fill": [_vp, _i64, _pi64, _i32, _i32, _i32, _i32, _i32, _i32, _i32, _i32, _vp, _vp, _i64, _pi64,
                       _pi64, _vp],
    "a3d_flood_extract": [_vp, _i32, _i32, _i32, _i32, _pi64, _vp, _vp, _i64, _vp],
    "a3d_flood_dense_mask": [_vp, _i32, _i32, _i32, _vp, _vp],
    "a3d_flood_dam_mask": [_vp, _i32, _i32, _i32, _vp, _i64, _pi64, _vp, _vp],
    "a3d_ijk_bounds": [_vp, _i64, _vp, _vp],
    "a3d_dmc_ws_bytes": [_i64, _pi64],
    "a3d_dmc_topology": [_vp, _vp, _i32, _f64, _vp, _i64, _i64, _i32, _vp, _vp, _i64, _pi64, _vp],
    "a3d_dmc_faces": [_vp, _i64, _pi64, _vp, _vp],
    "a3d_dmc_forward": [_vp, _i64, _pi64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _i32, _f64, _f64, _i32, _vp,
                        _vp, _vp, _vp, _vp, _vp],
    "a3d_dmc_backward": [_vp, _i64, _i64, _pi64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _vp, _i32, _f64, _f64,
                         _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp],
    "a3d_dmc_bwd_scratch_bytes": [_pi64, _i32, _pi64],
    "a3d_dmc_topology_get": [_vp, _i64, _pi64, _i32, _vp, _vp],
    "a3d_dmc_corner_grads": [_vp, _i64, _i64, _vp, _vp, _vp, _vp, _i32, _f64, _i32, _vp, _vp, _vp],
    "a3d_dmc_forward_qef": [_vp, _i64, _pi64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _i32, _f64, _f64, _i32,
                            _vp, _f64, _vp, _vp, _vp, _vp, _vp, _vp],
    "a3d_vox_ws_bytes": [_i64, _pi64],
    "a3d_vox_cells": [_vp, _pf32, _pf32, _pi64, _pi64, _vp, _i64, _pi64, _vp],
    "a3d_vox_cells_get": [_vp, _pi64, _pi64, _i64, _vp, _vp],
    "a3d_cells_sat_count": [_vp, _vp, _i64, _pf32, _pf32, _vp, _vp],
    "a3d_cells_sat_pairs": [_vp, _vp, _i64, _pf32, _pf32, _vp, _vp, _vp, _vp],
    "a3d_cells_ws_bytes": [_i64, _i32, _i32, _pi64],
    "a3d_unique_cells": [_vp, _i64, _i32, _i32, _vp, _i64, _vp, _pi64, _vp],
    "a3d_cell_vertices": [_vp, _i64, _i32, _i32, _vp, _i64, _vp, _pi64, _vp],
    "a3d_cell_vertices_get": [_vp, _i64, _i32, _i32, _i64, _vp, _vp, _vp, _pf32, _pf32, _vp],
    "a3d_cell_neighbors": [_vp, _i64, _i32, _i32, _vp, _i64, _vp, _vp],
    "a3d_wet_ws_bytes": [_i32, _pi64],
    "a3d_wet_mask": [_i32, _vp, _i64, _vp, _i64, _i32, _vp, _i64, _vp, _vp],
    "a3d_sg_sizes": [_pi64, _pi64],
    "a3d_sg_init": [_vp, _vp, _pf32, _pf32, _u64, _vp],
    "a3d_sg_sample": [_vp, _vp, _i64, _i64, _f32, _f32, _pf32, _pf32, _u64, _vp, _vp],
    "a3d_sg_grad": [_vp, _vp, _vp, _vp, _i64, _vp, _vp],
    "a3d_sg_table_grad": [_vp, _vp, _vp],
    "a3d_sg_apply": [_vp, _vp, _f32, _f32, _f32, _f32, _i64, _vp],
    "a3d_sg_infer": [_vp, _vp, _i64, _vp, _vp, _vp],
    "a3d_sg_route_ws_bytes": [_i64, _pi64],
    "a3d_sg_udf": [_vp, _vp, _vp, _i64, _vp, _f32, _vp, _vp, _vp, _vp, _vp, _i64, _pi64, _vp],
    "a3d_batch_order_ws_bytes": [_i64, _pi64],
    "a3d_batch_order": [_vp, _i32, _vp, _i64, _vp, _vp, _vp, _i64, _vp],
    "a3d_batch_udf": [_vp, _vp, _vp, _vp, _i64, _vp, _i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp],
    "a3d_graph_begin": [_vp],
    "a3d_graph_end": [_vp, _pi64],
    "a3d_graph_launch": [_i64, _vp],
    "a3d_graph_destroy": [_i64],
    "a3d_calib_valu": [_vp, _i64, _i32, _vp],
}


def _load():
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"atom3d_amd: HIP library not found at {LIB_PATH}; build it with `python -m atom3d_amd.build`"
            )
        # torch first: the library then binds to the HIP runtime torch already
        # loaded (same soname), so pointers and streams are shared.
        torch.cuda.is_available  # noqa: B018  (import side effect only)
        lib = ctypes.CDLL(LIB_PATH)
        for name, argtypes in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = argtypes
            fn.restype = ctypes.c_int
        lib.a3d_last_error.restype = ctypes.c_char_p
        lib.a3d_last_error.argtypes = []
        lib.a3d_version.restype = ctypes.c_char_p
        lib.a3d_version.argtypes = []
        _lib = lib
        return lib


def lib():
    return _lib if _lib is not None else _load()


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().a3d_last_error().decode(errors="replace")
        raise RuntimeError(f"atom3d_amd {what} failed (code {rc}): {msg}")


_roctx_fns = None  # (push, pop) when ATOM3D_ROCTX is set, else False; None until first asked


def _roctx():
    """roctx range functions when the environment variable ATOM3D_ROCTX is set
    to anything but "" or "0" (then every C-ABI call below runs inside a
    range named after the entry point, which `rocprofv3 --marker-trace`
    records), else False.  Off by default: two extra foreign calls per entry
    point cost about a microsecond."""
    global _roctx_fns
    if _roctx_fns is None:
        fns = False
        if os.environ.get("ATOM3D_ROCTX", "0") not in ("", "0"):
            try:
                r = ctypes.CDLL("libroctx64.so.4")
            except OSError:
                r = ctypes.CDLL("/opt/rocm/lib/libroctx64.so.4")
            r.roctxRangePushA.argtypes = [ctypes.c_char_p]
            r.roctxRangePushA.restype = ctypes.c_int
            r.roctxRangePop.argtypes = []
            r.roctxRangePop.restype = ctypes.c_int
            fns = (r.roctxRangePushA, r.roctxRangePop)
        _roctx_fns = fns
    return _roctx_fns


class trace_range:
    """``with trace_range("MeshBVH.udf"): ...`` — a roctx range around a block
    of host code when ATOM3D_ROCTX is set; nothing otherwise."""

    def __init__(self, name: str):
        self.name = name.encode()

    def __enter__(self):
        tr = _roctx()
        if tr:
            tr[0](self.name)
        return self

    def __exit__(self, *exc):
        tr = _roctx()
        if tr:
            tr[1]()
        return False


def call(name: str, *args) -> None:
    fn = getattr(lib(), name)
    tr = _roctx_fns if _roctx_fns is not None else _roctx()
    if not tr:
        check(fn(*args), name)
        return
    tr[0](name.encode())
    try:
        rc = fn(*args)
    finally:
        tr[1]()
    check(rc, name)


def ptr(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    return t.data_ptr()


def stream(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def require_cuda(t: torch.Tensor, name: str) -> None:
    if not t.is_cuda:
        raise RuntimeError(f"atom3d_amd: {name} must be a GPU tensor (got device {t.device})")


def version() -> str:
    return lib().a3d_version().decode()

"""Host-side checks (no GPU): the C-ABI library loads and exports every
declared symbol, argument validation errors surface as messages, the cube
tables equal the reference's, and the synthetic workloads are deterministic."""
import ctypes
import hashlib
import os
import re

import numpy as np
import pytest
import torch

from atom3d_amd import constants as C
from atom3d_amd import datasets
from atom3d_amd.kernels import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "atom3d_amd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(a3d_\w+)\s*\(", src, re.M)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), f"{n} declared in include/atom3d_amd.h but not exported"
    # and every binding the Python side sets up is declared
    assert set(_lib.SIGNATURES) <= set(names)


def _declarations():
    """name -> parameter type strings of every int-returning entry point in
    include/atom3d_amd.h."""
    src = open(os.path.join(ROOT, "include", "atom3d_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", " ", src, flags=re.S)
    out = {}
    for m in re.finditer(r"^\s*int\s+(a3d_\w+)\s*\(([^;]*?)\)\s*;", src, re.M | re.S):
        params = " ".join(m.group(2).split())
        out[m.group(1)] = [] if params in ("", "void") else [p.strip() for p in params.split(",")]
    return out


def _ctype_ok(param: str, t) -> bool:
    if "*" in param or param.startswith("a3d_stream_t"):
        return t is ctypes.c_void_p or (isinstance(t, type) and issubclass(t, ctypes._Pointer))
    base = param.rsplit(" ", 1)[0].replace("const ", "").strip()
    want = {"int64_t": ctypes.c_int64, "uint64_t": ctypes.c_uint64, "int": ctypes.c_int, "int32_t": ctypes.c_int32,
            "float": ctypes.c_float, "double": ctypes.c_double}.get(base)
    return want is not None and t is want


def test_every_entry_point_has_a_matching_binding():
    # a call without argtypes would pass 64-bit pointers and sizes as 32-bit C ints
    decl = _declarations()
    assert len(decl) >= 15
    missing = sorted(set(decl) - set(_lib.SIGNATURES))
    assert not missing, f"entry points without a ctypes signature in kernels/_lib.py: {missing}"
    for name, params in decl.items():
        sig = _lib.SIGNATURES[name]
        assert len(sig) == len(params), f"{name}: header has {len(params)} parameters, binding {len(sig)}"
        for i, (p, t) in enumerate(zip(params, sig)):
            assert _ctype_ok(p, t), f"{name} parameter {i} ({p!r}) bound as {t}"


def test_version_and_argument_errors():
    assert "gfx950" in _lib.version()
    b, w = ctypes.c_int64(), ctypes.c_int64()
    with pytest.raises(RuntimeError, match="leaf_max"):
        _lib.call("a3d_bvh_sizes", 100, 0, ctypes.byref(b), ctypes.byref(w))
    with pytest.raises(RuntimeError, match="null BVH blob"):
        _lib.call("a3d_bvh_udf", None, None, 5, None, None, None, None, None, 0, None, None)
    with pytest.raises(RuntimeError, match="unknown flags"):
        _lib.call("a3d_bvh_udf", 1, None, 5, None, None, None, None, None, 32, None, None)
    with pytest.raises(RuntimeError, match="level"):
        _lib.call("a3d_octree_level", 1, (ctypes.c_float * 6)(), 22, None, 0, 1, 1, 1, 1 << 20, None)
    with pytest.raises(RuntimeError, match="bad grid"):
        _lib.call("a3d_udf_grid_seed", None, None, None, None, 0, (ctypes.c_float * 3)(), 0.1, None, None, 5, 0.0,
                  None, None)
    with pytest.raises(RuntimeError, match="null argument"):
        _lib.call("a3d_udf_grid_seed", 1, 1, None, None, 8, (ctypes.c_float * 3)(), 0.1, 1, None, 5, 0.0, 1, None)
    with pytest.raises(RuntimeError, match="bad grid"):
        _lib.call("a3d_bvh_udf_grid", 1, 1, 5, None, 1, None, None, None, 8, (ctypes.c_float * 3)(), -1.0, 0.0,
                  1, 1, 1, None, 0, None)
    with pytest.raises(RuntimeError, match="split walk"):
        _lib.call("a3d_bvh_udf_grid", 1, 1, 5, None, 1, None, None, None, 8, (ctypes.c_float * 3)(), 0.1, 0.0,
                  1, 1, 1, None, 8, None)
    _lib.call("a3d_bvh_sizes", 4096000, 4, ctypes.byref(b), ctypes.byref(w))
    assert b.value == 256 + 64 * 4095999 + (48 + 32) * 4096000 + 96 * 4095999 + 48 * 4096000


def test_cube_tables_match_reference_values():
    # literal values of atom3d/constants.py:11-64 (fixture data)
    assert C.CUBE_CORNERS.tolist() == [[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0],
                                       [0, 0, 1], [1, 0, 1], [0, 1, 1], [1, 1, 1]]
    assert C.CUBE_EDGES.tolist() == [[0, 1], [2, 3], [4, 5], [6, 7], [0, 2], [1, 3], [4, 6], [5, 7],
                                     [0, 4], [1, 5], [2, 6], [3, 7]]
    assert C.NEIGHBOR_OFFSETS_26.shape == (26, 3)
    assert C.NEIGHBOR_OFFSETS_26[0].tolist() == [-1, -1, -1] and C.NEIGHBOR_OFFSETS_26[13].tolist() == [1, 0, 0]
    assert C.EDGE_NEIGHBOR_OFFSETS[0].tolist() == [[0, 1, 0], [0, 0, 1], [0, 1, 1]]
    assert C.EDGE_NEIGHBOR_OFFSETS[3].tolist() == [[0, -1, 0], [0, 0, -1], [0, -1, -1]]
    assert C.EDGE_NEIGHBOR_OFFSETS[6].tolist() == [[1, 0, 0], [0, 0, -1], [1, 0, -1]]
    assert C.EDGE_NEIGHBOR_OFFSETS[11].tolist() == [[-1, 0, 0], [0, -1, 0], [-1, -1, 0]]
    assert C.CORNER_NEIGHBOR_OFFSETS[7].tolist()[-1] == [1, 1, 1]
    assert C.CORNER_NEIGHBOR_OFFSETS[0].tolist()[-1] == [-1, -1, -1]


def test_icosphere_counts_orientation_and_determinism():
    v, f = datasets.icosphere(5)
    assert v.shape == (10242, 3) and f.shape == (20480, 3)
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    assert np.all(np.einsum("ij,ij->i", np.cross(b - a, c - a), a + b + c) > 0)  # outward winding
    V, F = datasets.noisy_icosphere()
    V2, F2 = datasets.noisy_icosphere()
    assert torch.equal(V, V2) and torch.equal(F, F2)
    assert V.dtype == torch.float32 and F.dtype == torch.int32
    Vu, Fu = datasets.sphere_union(n_spheres=3)
    assert Fu.shape == (3 * 20480, 3) and int(Fu.max()) == 3 * 10242 - 1


def test_octree_cpu_geometry_helpers():
    from atom3d_amd import OctreeIndexer
    oc = OctreeIndexer(max_level=4, device="cpu")
    cubes = oc.all_cubes_at_level(2)
    assert cubes.shape == (64, 3) and cubes[1].tolist() == [0, 0, 1]
    kids = oc.subdivide(cubes[:1] + 1, 2)
    assert kids.tolist()[:3] == [[2, 2, 2], [3, 2, 2], [2, 3, 2]]
    lo, hi = oc.cube_aabb_level(torch.tensor([[1, 2, 3]]), 2)
    assert lo.tolist() == [[-0.5, 0.0, 0.5]] and hi.tolist() == [[0.0, 0.5, 1.0]]
    idx = oc.ijk_to_cube_level(torch.tensor([[1, 2, 3]]), 2)
    assert idx.item() == (1 * 4 + 2) * 4 + 3
    assert oc.cube_to_ijk_level(idx, 2).tolist() == [[1, 2, 3]]


def test_new_entry_points_validate_arguments():
    """Argument errors of the DMC, surrogate, batch, graph and flood entry
    points come back as messages before any device work."""
    n = ctypes.c_int64()
    with pytest.raises(RuntimeError, match="out of range"):
        _lib.call("a3d_dmc_ws_bytes", -1, ctypes.byref(n))
    with pytest.raises(RuntimeError, match="dtype"):
        _lib.call("a3d_dmc_topology", None, None, 7, 0.0, None, 10, 10, 4, None, None, 0, None, None)
    sizes = (ctypes.c_int64 * 5)()
    with pytest.raises(RuntimeError, match="null"):
        _lib.call("a3d_dmc_forward", None, 1, sizes, None, None, None, None, None, None, None, 0, 0, 0.0, 0.99, 4,
                  None, None, None, None, None, None)
    with pytest.raises(RuntimeError, match="dtype"):
        _lib.call("a3d_dmc_corner_grads", 1, 1, 8, None, None, None, None, 3, 0.0, 4, None, None, None)
    with pytest.raises(RuntimeError, match="null argument"):
        _lib.call("a3d_dmc_corner_grads", 1, 1, 8, None, None, None, None, 0, 0.0, 4, None, None, None)
    with pytest.raises(RuntimeError, match="resolution"):
        _lib.call("a3d_dmc_corner_grads", 1, 1, 8, None, None, None, None, 0, 0.0, 0, None, None, None)
    sizes[1] = 1
    with pytest.raises(RuntimeError, match="null argument"):
        _lib.call("a3d_dmc_forward_qef", 1, 1, sizes, None, None, None, None, None, None, None, 0, 0, 0.0, 0.99, 4,
                  None, 0.01, None, None, None, None, None, None)
    with pytest.raises(RuntimeError, match="qef_reg_scale"):
        _lib.call("a3d_dmc_forward_qef", 1, 1, sizes, None, None, None, None, None, None, None, 0, 0, 0.0, 0.99, 4,
                  None, -1.0, None, None, None, None, None, None)
    with pytest.raises(RuntimeError, match="bad arguments"):
        _lib.call("a3d_cells_ws_bytes", -1, 8, 0, ctypes.byref(n))
    with pytest.raises(RuntimeError, match="bitmap method"):
        _lib.call("a3d_cells_ws_bytes", 10, 4096, 2, ctypes.byref(n))
    _lib.call("a3d_cells_ws_bytes", 10, 1024, 0, ctypes.byref(n))
    small = n.value
    _lib.call("a3d_cells_ws_bytes", 10, 1024, 2, ctypes.byref(n))
    assert n.value > 1024 ** 3 // 8 > small  # 10 cells on 1024^3: the sort path
    with pytest.raises(RuntimeError, match="null argument"):
        _lib.call("a3d_unique_cells", None, 4, 8, 0, None, 0, None, ctypes.byref(n), None)
    with pytest.raises(RuntimeError, match="bad arguments"):
        _lib.call("a3d_cell_vertices", None, 4, 0, 0, None, 0, None, ctypes.byref(n), None)
    with pytest.raises(RuntimeError, match="null argument"):
        _lib.call("a3d_cell_neighbors", None, 4, 8, 1, None, 0, None, None)
    with pytest.raises(RuntimeError, match="kind"):
        _lib.call("a3d_wet_mask", 3, None, 1, None, 0, 8, None, 0, None, None)
    with pytest.raises(RuntimeError, match="bitmap"):
        _lib.call("a3d_wet_mask", 0, 1, 1, None, 0, 8, 1, 8, 1, None)
    _lib.call("a3d_wet_ws_bytes", 1024, ctypes.byref(n))
    assert n.value == 1024 ** 3 // 8
    with pytest.raises(RuntimeError, match="null"):
        _lib.call("a3d_sg_init", None, None, None, None, 0, None)
    with pytest.raises(RuntimeError, match="bad batch"):
        _lib.call("a3d_sg_grad", 1, 1, None, None, 0, None, None)
    with pytest.raises(RuntimeError, match="launch shape"):
        _lib.call("a3d_batch_udf", None, None, None, None, 5, None, 0, 1, None, None, None, None, None, None)
    with pytest.raises(RuntimeError, match="null handle"):
        _lib.call("a3d_graph_launch", 0, None)
    with pytest.raises(RuntimeError, match="bad arguments"):
        _lib.call("a3d_ijk_bounds", None, 5, None, None)
    with pytest.raises(RuntimeError, match="null seed"):
        _lib.call("a3d_bvh_udf_seeded", 1, None, 5, None, None, None, None, None, None, 0, None)
    with pytest.raises(RuntimeError, match="null work"):
        _lib.call("a3d_bvh_udf_work", 1, 1, 5, None, 1, 1, 1, None, 0, None, None)
    with pytest.raises(RuntimeError, match="need the wave walk"):
        _lib.call("a3d_bvh_udf_work", 1, 1, 5, None, 1, 1, 1, None, 1, 1, None)
    with pytest.raises(RuntimeError, match="not in"):
        _lib.call("a3d_query_order_work_ws_bytes", -1, ctypes.byref(n))
    _lib.call("a3d_query_order_work_ws_bytes", 1 << 20, ctypes.byref(n))
    plain = ctypes.c_int64()
    _lib.call("a3d_query_order_ws_bytes", 1 << 20, ctypes.byref(plain))
    assert n.value >= plain.value + 4 * (1 << 20) + 16 * (1 << 14)  # + the Morton order and 4 packet arrays
    lo3 = (ctypes.c_float * 3)()
    with pytest.raises(RuntimeError, match="bad grid"):
        _lib.call("a3d_query_order_work", 1, 5, 1, 0, lo3, 0.1, 1, 1, 1 << 20, None)
    with pytest.raises(RuntimeError, match="null argument"):
        _lib.call("a3d_query_order_work", 1, 5, None, 8, lo3, 0.1, 1, 1, 1 << 20, None)
    with pytest.raises(RuntimeError, match="workspace"):
        _lib.call("a3d_query_order_work", 1, 5, 1, 8, lo3, 0.1, 1, 256, 1, None)
    pb, sb = ctypes.c_int64(), ctypes.c_int64()
    _lib.call("a3d_sg_sizes", ctypes.byref(pb), ctypes.byref(sb))
    assert pb.value == 512 + 16 * 2 ** 19 * 4 + 64 * 32 * 2 + 64 * 64 * 2 + 64 * 4
    assert sb.value > 4 * 16 * 2 ** 19 * 8

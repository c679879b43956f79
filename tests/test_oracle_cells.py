"""The glue oracle (oracle/cells.py) pinned on the CPU: the neighbour tables
by the digest of the reference's literal tables, every function by brute
force over Python sets and dicts on small inputs, the unchecked-flat-key
neighbour lookup by hand-computed cases."""
import hashlib
import itertools
import json
import os

import numpy as np
import pytest

from atom3d_amd import constants as C
from oracle import cells as oc

DIGEST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "constants_digest.json")


def test_neighbor_tables_match_reference_digest():
    d = json.load(open(DIGEST))
    for k in ("NEIGHBOR_OFFSETS_6", "EDGE_NEIGHBOR_OFFSETS"):
        a = np.ascontiguousarray(getattr(C, k).numpy().astype(np.int32))
        assert list(a.shape) == d[k]["shape"]
        assert hashlib.sha256(a.tobytes()).hexdigest() == d[k]["sha256"], k
    # corners: 7 sign combinations of each corner's (2b - 1) per axis
    for i in range(8):
        s = [2 * ((i >> a) & 1) - 1 for a in range(3)]
        want = {tuple(s[a] if (m >> a) & 1 else 0 for a in range(3)) for m in range(1, 8)}
        assert {tuple(r) for r in C.CORNER_NEIGHBOR_OFFSETS[i].tolist()} == want


def _random_cells(rng, R, n):
    return rng.integers(0, R, (n, 3))


def test_unique_cells_and_vertices_brute_force():
    rng = np.random.default_rng(0)
    R = 9
    cells = _random_cells(rng, R, 300)  # with duplicates
    u = oc.unique_cells(cells)
    assert [tuple(r) for r in u] == sorted({tuple(r) for r in cells})
    v, cube = oc.cell_vertices(u)
    corners = sorted({(a + dx, b + dy, c + dz) for a, b, c in u.tolist()
                      for dz in (0, 1) for dy in (0, 1) for dx in (0, 1)})
    assert [tuple(r) for r in v] == corners
    pos = {p: i for i, p in enumerate(corners)}
    for n, (a, b, c) in enumerate(u.tolist()):
        for k in range(8):
            assert cube[n, k] == pos[(a + (k & 1), b + ((k >> 1) & 1), c + ((k >> 2) & 1))]
    w = oc.world(v, [-1.0, -1.0, -1.0], np.float32(2.0) / np.float32(R))
    assert w.dtype == np.float32
    np.testing.assert_allclose(w, v * (2.0 / R) - 1.0, rtol=0, atol=1e-6)


def test_neighbors_brute_force_and_flat_key_wrap():
    rng = np.random.default_rng(1)
    R = 6
    u = oc.unique_cells(_random_cells(rng, R, 150))
    adj = oc.cell_neighbors(u, R)
    index = {tuple(r): i for i, r in enumerate(u.tolist())}
    dirs = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    for n, (a, b, c) in enumerate(u.tolist()):
        for d, (dx, dy, dz) in enumerate(dirs):
            x, y, z = a + dx, b + dy, c + dz
            if 0 <= x < R and 0 <= y < R and 0 <= z < R:
                assert adj[n, d] == index.get((x, y, z), -1)
    # off the +y face of cell (0, R-1, 2) the flat key R^2 + 2 is cell (1, 0, 2)
    cells = np.array([[0, R - 1, 2], [1, 0, 2], [2, 3, 0], [2, 2, R - 1]])
    u = oc.unique_cells(cells)
    adj = oc.cell_neighbors(u, R)
    row = {tuple(r): i for i, r in enumerate(u.tolist())}
    assert adj[row[(0, R - 1, 2)], 2] == row[(1, 0, 2)]
    assert adj[row[(1, 0, 2)], 3] == row[(0, R - 1, 2)]
    # off the -z face of (2, 3, 0) lands on (2, 2, R-1); off +z of that, back
    assert adj[row[(2, 3, 0)], 5] == row[(2, 2, R - 1)]
    assert adj[row[(2, 2, R - 1)], 4] == row[(2, 3, 0)]
    # off the -x face of (0, ...) is a negative key: none
    assert adj[row[(0, R - 1, 2)], 1] == -1


@pytest.mark.parametrize("kind", ["faces", "edges", "corners"])
def test_wet_masks_brute_force(kind):
    rng = np.random.default_rng(2)
    R = 7
    dam = oc.unique_cells(_random_cells(rng, R, 80))
    coll = oc.unique_cells(_random_cells(rng, R, 120))
    got = getattr(oc, f"wet_{kind}")(dam, coll, R)
    table = {"faces": C.NEIGHBOR_OFFSETS_6.numpy()[:, None, :], "edges": C.EDGE_NEIGHBOR_OFFSETS.numpy(),
             "corners": C.CORNER_NEIGHBOR_OFFSETS.numpy()}[kind]
    cs = {tuple(r) for r in coll.tolist()}
    for n, p in enumerate(dam.tolist()):
        for k in range(table.shape[0]):
            want = any(tuple(int(p[a] + o[a]) for a in range(3)) in cs for o in table[k]
                       if all(0 <= p[a] + o[a] < R for a in range(3)))
            assert got[n, k] == want, (p, k)


def test_dual_graph_of_one_cell_and_a_full_block():
    lo, hi = [-1.0, -1.0, -1.0], [1.0, 1.0, 1.0]
    w, v, cube, adj, cells = oc.construct_surface_dual_graph(np.array([[3, 4, 5]]), np.zeros((0, 3)), None, 8, lo, hi)
    # vertices are in lexicographic (x-major) order, corners x-fastest
    assert cells.tolist() == [[3, 4, 5]] and len(v) == 8 and cube.tolist() == [[0, 4, 2, 6, 1, 5, 3, 7]]
    assert (adj == -1).all()
    # a 2x2x2 block: 27 vertices, every interior face neighbour present
    blk = np.array(list(itertools.product(range(2, 4), repeat=3)))
    w, v, cube, adj, cells = oc.construct_surface_dual_graph(blk[:3], blk[3:6], blk[6:], 8, lo, hi)
    assert len(cells) == 8 and len(v) == 27
    assert (adj >= 0).sum() == 8 * 3

"""Dual-marching-cubes case tables, generated from their definition.

The dual vertices of a cube are the connected components of its zero
crossing: edges with a sign change (inside = sdf < 0) are joined when they
bound the same iso-line segment on a cube face; on an ambiguous face (two
diagonal inside corners) the inside corners are kept apart.  Each component
lists its edges in ascending order and the components are ordered by the
smallest corner on the minority side (inside corners when at most four are
inside, outside corners otherwise).  A case needs a neighbour check when it
has one component, its complement has two, and it has exactly one ambiguous
face: if the cube across that face is also such a case, the complement's
grouping is used.

These definitions reproduce, entry for entry, the FlexiCubes tables the
reference ships (Atom3D atom3d/mesh_extractor/tables.py: dmc_table,
num_vd_table, check_table); tests/test_dmc_tables.py pins them against a
digest of the reference's values.  Edge numbering follows the reference's
cube_edges (sparse_diffdmc.py:73-77), corners CUBE_CORNERS.
"""
from __future__ import annotations

import numpy as np

# local edge e joins corners EDGE_CORNERS[e]
EDGE_CORNERS = ((0, 1), (1, 5), (4, 5), (0, 4), (2, 3), (3, 7), (6, 7), (2, 6), (2, 0), (3, 1), (7, 5), (6, 4))


def _corner(i):
    return (i & 1, (i >> 1) & 1, (i >> 2) & 1)


def _faces():
    """Six faces as (axis, side, corners in cyclic order)."""
    out = []
    ring = {(0, 0): 0, (1, 0): 1, (1, 1): 2, (0, 1): 3}
    for axis in range(3):
        o = [k for k in range(3) if k != axis]
        for side in (0, 1):
            cs = [i for i in range(8) if _corner(i)[axis] == side]
            cs.sort(key=lambda i: ring[(_corner(i)[o[0]], _corner(i)[o[1]])])
            out.append((axis, side, cs))
    return out


_FACES = _faces()
_EDGE_ID = {frozenset(p): e for e, p in enumerate(EDGE_CORNERS)}


def _groups(case: int):
    inside = [(case >> i) & 1 for i in range(8)]
    cross = [e for e, (a, b) in enumerate(EDGE_CORNERS) if inside[a] != inside[b]]
    parent = {e: e for e in cross}

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    def join(a, b):
        parent[find(a)] = find(b)

    for _, _, cs in _FACES:
        ring = [_EDGE_ID[frozenset((cs[i], cs[(i + 1) % 4]))] for i in range(4)]
        hot = [e for e in ring if e in parent]
        if len(hot) == 2:
            join(hot[0], hot[1])
        elif len(hot) == 4:  # ambiguous face: cut off each inside corner
            for i in range(4):
                if inside[cs[i]]:
                    join(ring[i - 1], ring[i])
    comps = {}
    for e in cross:
        comps.setdefault(find(e), []).append(e)
    minority = 1 if sum(inside) <= 4 else 0

    def key(g):
        return min(a if inside[a] == minority else b for a, b in (EDGE_CORNERS[e] for e in g))

    return sorted((sorted(g) for g in comps.values()), key=key)


def _ambiguous_faces(case: int):
    inside = [(case >> i) & 1 for i in range(8)]
    out = []
    for axis, side, cs in _FACES:
        s = [inside[c] for c in cs]
        if s[0] == s[2] and s[1] == s[3] and s[0] != s[1]:
            d = [0, 0, 0]
            d[axis] = 1 if side else -1
            out.append(d)
    return out


def build_tables():
    dmc = -np.ones((256, 4, 7), np.int64)
    num_vd = np.zeros(256, np.int64)
    check = np.zeros((256, 5), np.int64)
    for c in range(256):
        gs = _groups(c)
        num_vd[c] = len(gs)
        for k, g in enumerate(gs):
            dmc[c, k, :len(g)] = g
    for c in range(256):
        amb = _ambiguous_faces(c)
        if num_vd[c] == 1 and num_vd[255 - c] == 2 and len(amb) == 1:
            check[c] = [1, *amb[0], 255 - c]
    return dmc, num_vd, check


dmc_table, num_vd_table, check_table = build_tables()

__all__ = ["EDGE_CORNERS", "dmc_table", "num_vd_table", "check_table", "build_tables"]

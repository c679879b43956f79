"""Cube topology tables (x-first corner order).

The values are part of the drop-in contract — octree child order, flood-fill
neighbourhoods and the DMC tables index them — so they equal Atom3D
atom3d/constants.py:11-111 entry for entry.  They are generated here from
their definitions rather than listed.
"""
import torch

# Corner i sits at ((i >> 0) & 1, (i >> 1) & 1, (i >> 2) & 1).
CUBE_CORNERS = torch.tensor([[(i >> 0) & 1, (i >> 1) & 1, (i >> 2) & 1] for i in range(8)], dtype=torch.int64)

# 12 edges as corner pairs: 4 x-aligned, 4 y-aligned, 4 z-aligned; within an
# axis the edges follow the corner order of their lower endpoint.
def _edges():
    out = []
    for axis in range(3):
        bit = 1 << axis
        lows = [c for c in range(8) if not c & bit]
        out += [[c, c | bit] for c in lows]
    return out


CUBE_EDGES = torch.tensor(_edges(), dtype=torch.int64)

# 6 faces (4 corners each, counter-clockwise seen from outside): -X, +X, -Y, +Y, -Z, +Z.
CUBE_FACES = torch.tensor([
    [0, 4, 6, 2], [1, 3, 7, 5],
    [0, 1, 5, 4], [2, 6, 7, 3],
    [0, 2, 3, 1], [4, 5, 7, 6],
], dtype=torch.int64)

EDGE_TO_FACES = torch.tensor([
    [0, 2], [0, 5], [0, 3], [0, 4],
    [1, 2], [1, 5], [1, 3], [1, 4],
    [2, 4], [2, 5], [3, 5], [3, 4],
], dtype=torch.int64)

CUBE_CORNER_SIGNS = (CUBE_CORNERS * 2 - 1).float()

# 6-neighbourhood: +x, -x, +y, -y, +z, -z.
NEIGHBOR_OFFSETS_6 = torch.tensor(
    [[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], dtype=torch.int32)

# 26-neighbourhood, z-major then y then x, centre excluded.
NEIGHBOR_OFFSETS_26 = torch.tensor(
    [[dx, dy, dz] for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1) if (dx, dy, dz) != (0, 0, 0)],
    dtype=torch.int32)


def _edge_neighbors():
    # For an edge along `axis` at local position (u, w) on the two other axes,
    # the voxels sharing it are offset by -/+1 along those axes (toward the
    # edge's side), and both.
    out = []
    for e in range(12):
        c0, c1 = CUBE_EDGES[e].tolist()
        axis = [k for k in range(3) if (c0 ^ c1) >> k & 1][0]
        others = [k for k in range(3) if k != axis]
        s = [1 if not (c0 >> k) & 1 else -1 for k in range(3)]
        o1 = [0, 0, 0]; o1[others[0]] = s[others[0]]
        o2 = [0, 0, 0]; o2[others[1]] = s[others[1]]
        o3 = [0, 0, 0]; o3[others[0]] = s[others[0]]; o3[others[1]] = s[others[1]]
        out.append([o1, o2, o3])
    return out


EDGE_NEIGHBOR_OFFSETS = torch.tensor(_edge_neighbors(), dtype=torch.int32)

CORNER_NEIGHBOR_OFFSETS = torch.empty((8, 7, 3), dtype=torch.int32)
for _i in range(8):
    _sx, _sy, _sz = CUBE_CORNER_SIGNS[_i].int().tolist()
    CORNER_NEIGHBOR_OFFSETS[_i] = torch.tensor([
        [_sx, 0, 0], [0, _sy, 0], [0, 0, _sz],
        [_sx, _sy, 0], [_sx, 0, _sz], [0, _sy, _sz],
        [_sx, _sy, _sz],
    ], dtype=torch.int32)

SIGN_WATER = 1
SIGN_DRY = -1
SIGN_DAM_DEFAULT = -1

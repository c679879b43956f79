"""The voxelizer oracle (oracle/voxelizer.py) pinned on the CPU: the
candidate block by its defining property (tight cover of every face's box,
clamped), the CSR by brute force (every (cell, face) pair tested with the
restated SAT, oracle.tri_aabb_sat, in Python loops)."""
import numpy as np

import oracle
from atom3d_amd import datasets
from oracle import voxelizer as ov


def _mesh():
    v, f = datasets.icosphere(1, 0.7)
    return v.astype(np.float32), f.astype(np.int32)


def test_candidate_block_covers_every_face_box_tightly():
    V, F = _mesh()
    res = 12
    b0, cell = np.float32([-1, -1, -1]), np.float32(2.0) / np.float32(res) * np.ones(3, np.float32)
    origin, dims = ov.candidate_block(V, F, b0, cell, res)
    t = V[F]
    for lo, hi in zip(t.min(1), t.max(1)):
        a = np.clip(np.floor((lo - b0) / cell), 0, res - 1)
        b = np.clip(np.ceil((hi - b0) / cell), 0, res - 1)
        assert (a >= origin).all() and (b <= origin + dims - 1).all()
    lo_all = np.clip(np.floor((t.min(1) - b0) / cell), 0, res - 1).min(0)
    hi_all = np.clip(np.ceil((t.max(1) - b0) / cell), 0, res - 1).max(0)
    assert (origin == lo_all).all() and (origin + dims - 1 == hi_all).all()


def test_csr_matches_brute_force_sat():
    V, F = _mesh()
    res = 6
    b0, cell = np.float32([-1, -1, -1]), np.full(3, np.float32(2.0) / np.float32(res))
    for strategy in ("candidate", "all"):
        vox, faces, start, count = ov.voxelize_with_faces(V, F, b0, cell, res, strategy)
        want = {}
        for i in range(res):
            for j in range(res):
                for k in range(res):
                    lo = np.float32([i, j, k]) * cell + b0
                    hi = lo + cell
                    hits = [fi for fi in range(len(F)) if oracle.tri_aabb_sat(V[F[fi]].reshape(-1), lo, hi)]
                    if hits:
                        want[(i, j, k)] = hits
        assert [tuple(r) for r in vox] == sorted(want)  # lexicographic = meshgrid('ij') order
        for r, key in enumerate(map(tuple, vox)):
            assert faces[start[r]:start[r] + count[r]].tolist() == want[key]

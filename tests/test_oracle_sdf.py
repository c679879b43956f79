"""The pseudo-normal sign oracle (oracle/sdf.py) pinned on the CPU by the
signs a closed convex mesh fixes: points at +-delta along the normal rays
through its vertices (vertex pseudo-normal region), edge midpoints (edge
region) and face centroids (face region), with closest points from the
oracle's brute-force closest-point search."""
import numpy as np

import oracle
from atom3d_amd import datasets
from oracle.sdf import compute_sign, pseudo_normals


def test_sign_oracle_on_normal_rays_of_icosphere():
    V, F = datasets.icosphere(2, radius=0.5)
    V = V.astype(np.float32)
    mids = (V[F[:, 0]] + V[F[:, 1]]) / 2
    cent = V[F].mean(1)
    base = np.concatenate([V, mids, cent])
    u = base / np.linalg.norm(base, axis=1, keepdims=True)
    P = np.concatenate([base + 1e-3 * u, base - 1e-3 * u]).astype(np.float32)
    expect = np.concatenate([np.ones(len(base)), -np.ones(len(base))])
    d, f, c, w = oracle.brute_udf(V, F, P)
    fn, vpn, epn, he = pseudo_normals(V, F)
    s, _ = compute_sign(P, c, f, w, fn, vpn, epn, he, F)
    assert np.array_equal(s, expect)
    # every region is exercised: vertex (two zero barycentrics), edge (one), face (none)
    nz = (w < 1e-4).sum(1)
    assert (nz == 2).sum() >= len(V) and (nz == 1).sum() >= len(mids) and (nz == 0).sum() >= len(cent)


def test_pseudo_normals_unit_and_outward():
    V, F = datasets.icosphere(3, radius=0.5)
    fn, vpn, epn, he = pseudo_normals(V, F)
    for n in (fn, vpn, epn):
        assert np.allclose(np.linalg.norm(n, axis=1), 1, atol=1e-6)
    assert np.all((vpn * V).sum(1) > 0)
    assert len(epn) == 3 * len(F) // 2 and he.max() == len(epn) - 1
    # no hit -> +1
    s, _ = compute_sign(np.zeros((2, 3)), np.zeros((2, 3)), np.array([-1, -1]), np.zeros((2, 3)), fn, vpn, epn, he, F)
    assert np.array_equal(s, [1, 1])

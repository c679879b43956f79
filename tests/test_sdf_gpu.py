"""MeshBVH.sdf (pseudo-normal sign, HIP kernel) against the oracle's
restatement of the reference's sign rule (oracle/sdf.py) and against signs
the geometry fixes on a closed convex mesh: points at +-delta along the
normal rays through vertices, edge midpoints and face centroids, and points
well inside or outside."""
import numpy as np
import pytest
import torch

from atom3d_amd import MeshBVH, datasets
from oracle.sdf import compute_sign, pseudo_normals

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _rays(V, F, delta):
    V = V.astype(np.float64)
    mids = (V[F[:, 0]] + V[F[:, 1]]) / 2
    cent = V[F].mean(1)
    base = np.concatenate([V, mids, cent])
    u = base / np.linalg.norm(base, axis=1, keepdims=True)
    # offset along the outward direction from the surface point
    out = base + delta * u
    inn = base - delta * u
    P = np.concatenate([out, inn]).astype(np.float32)
    expect = np.concatenate([np.ones(len(base)), -np.ones(len(base))]).astype(np.float32)
    return P, expect


@pytest.mark.parametrize("sub", [2, 4])
def test_sdf_signs_on_convex_mesh(sub):
    V, F = datasets.icosphere(sub, radius=0.5)
    P, expect = _rays(V, F, 1e-3)
    rng = np.random.default_rng(sub)
    d = rng.standard_normal((4000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.concatenate([rng.uniform(0.0, 0.4, 2000), rng.uniform(0.55, 1.5, 2000)])
    P = np.concatenate([P, (d * r[:, None]).astype(np.float32)])
    expect = np.concatenate([expect, np.where(r < 0.45, -1, 1).astype(np.float32)])
    m = MeshBVH(torch.tensor(V, dtype=torch.float32, device=DEV), torch.tensor(F, dtype=torch.int32, device=DEV))
    Pt = torch.tensor(P, device=DEV)
    res = m.sdf(Pt, return_closest=True, return_uvw=True, return_face_ids=True)
    s = torch.sign(res.distances).cpu().numpy()
    s[s == 0] = 1
    assert np.array_equal(s, expect)
    u = m.udf(Pt)
    assert torch.equal(res.distances.abs(), u)


def test_sdf_sign_matches_oracle_rule():
    V, F = datasets.noisy_icosphere()
    Vn, Fn = V.numpy(), F.numpy()
    m = MeshBVH(V.to(DEV), F.to(DEV))
    P = datasets.uniform_points(131072, seed=0)
    # plus points close to vertices and edges, where the vertex and edge
    # pseudo-normals are chosen
    rng = np.random.default_rng(1)
    pv = Vn[rng.integers(0, len(Vn), 20000)] * rng.uniform(0.97, 1.03, (20000, 1)).astype(np.float32)
    P = torch.cat([P, torch.tensor(pv)])
    res = m.sdf(P.to(DEV), return_closest=True, return_uvw=True, return_face_ids=True)
    fn, vpn, epn, he = pseudo_normals(Vn, Fn)
    signs, dots = compute_sign(P.numpy(), res.closest_points.cpu().numpy(), res.face_ids.cpu().numpy(),
                               res.uvw.cpu().numpy(), fn, vpn, epn, he, Fn)
    got = np.sign(res.distances.cpu().numpy())
    got[got == 0] = 1
    d = res.distances.abs().cpu().numpy()
    nz = res.uvw.cpu().numpy() < 1e-4
    assert (nz.sum(1) == 2).sum() > 1000 and (nz.sum(1) == 1).sum() > 1000
    # the pseudo-normals are computed by torch on the GPU and by numpy here:
    # a sign may differ only where the dot product is at rounding level
    bad = (got != signs) & (np.abs(dots) > 1e-6 * np.maximum(d, 1e-3))
    assert not bad.any(), np.nonzero(bad)[0][:10]
    assert (got != signs).mean() < 1e-4

"""The DMC oracle (oracle/dmc.py) pinned on the CPU: against the golden
vectors written from the reference's SparseDiffDMC (tests/golden/
dmc_fixtures.npz, made by tests/golden/make_dmc_fixtures.py), and against
what the geometry fixes independently (a plane's crossings lie on the plane,
a sphere's mesh is closed, oriented and of genus 0, the adjoint matches
central finite differences)."""
import os

import numpy as np
import pytest
import torch

from atom3d_amd.datasets import shapes_sdf
from oracle.dmc import dmc_forward, grid_band

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dmc_fixtures.npz")
CASES = ["shapes16", "random6", "random6_sparse"]
OUTS = ["vertices", "L_dev", "p_alpha", "n_alpha", "v_features"]
GRADS = ["sdf", "alpha", "beta", "deform", "v_features"]


@pytest.fixture(scope="module")
def fx():
    return dict(np.load(FIX))


def _inputs(fx, case, dt, full, requires_grad=False):
    kw = {}
    names = ["sdf"] + (["alpha", "beta", "deform", "v_features"] if full else [])
    for k in names:
        kw[k] = torch.tensor(fx[f"{case}/{k}"], dtype=dt, requires_grad=requires_grad)
    return torch.tensor(fx[f"{case}/voxel_coords"]).long(), torch.tensor(fx[f"{case}/cube_idx"]).long(), \
        int(fx[f"{case}/res"]), kw


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("tag", ["f64", "f32"])
@pytest.mark.parametrize("variant", ["plain", "full"])
def test_oracle_matches_reference_vectors(fx, case, tag, variant):
    dt = torch.float64 if tag == "f64" else torch.float32
    key = f"{case}_{tag}_{variant}"
    vc, ci, res, kw = _inputs(fx, case, dt, variant == "full", requires_grad=True)
    sdf = kw.pop("sdf")
    o = dmc_forward(vc, sdf, ci, res, dtype=dt, **kw)
    assert np.array_equal(o["faces"].numpy(), fx[f"{key}/out_faces"])
    atol = 1e-12 if tag == "f64" else 2e-6
    loss = 0
    for k in OUTS:
        if f"{key}/out_{k}" not in fx:
            assert o[k] is None
            continue
        np.testing.assert_allclose(o[k].detach().numpy(), fx[f"{key}/out_{k}"], rtol=0, atol=atol, err_msg=k)
        loss = loss + (torch.tensor(fx[f"{key}/W_{k}"], dtype=dt) * o[k]).sum()
    loss.backward()
    leaves = {"sdf": sdf, **kw}
    for k in GRADS:
        if f"{key}/grad_{k}" not in fx:
            continue
        ref = fx[f"{key}/grad_{k}"]
        scale = max(1.0, np.abs(ref).max())
        tol = 1e-10 if tag == "f64" else 2e-4
        np.testing.assert_allclose(leaves[k].grad.numpy(), ref, rtol=0, atol=tol * scale, err_msg=k)


@pytest.mark.parametrize("case", CASES)
def test_oracle_eval_split_matches_reference(fx, case):
    vc, ci, res, kw = _inputs(fx, case, torch.float64, False)
    o = dmc_forward(vc, kw["sdf"], ci, res, dtype=torch.float64, training=False)
    key = f"{case}_f64_plain"
    assert np.array_equal(o["faces"].numpy(), fx[f"{key}/eval_faces"])
    np.testing.assert_allclose(o["vertices"].numpy(), fx[f"{key}/eval_vertices"], rtol=0, atol=1e-12)


def test_plane_crossings_and_dual_vertices_lie_on_the_plane():
    n = np.array([0.3, -0.5, 0.81])
    n /= np.linalg.norm(n)
    cubes, s, ci = grid_band(lambda P: P @ n - 0.1, 12, 0.5)
    o = dmc_forward(torch.tensor(cubes), torch.tensor(s), torch.tensor(ci), 12)
    v = o["vertices"].numpy()
    assert len(v) > 50
    # vd = sum(beta u) / (sum(beta) + 1e-8): the 1e-8 shrinks vd by ~1e-8 / k
    assert np.abs(v @ n - 0.1).max() < 1e-8
    assert np.abs(o["p_alpha"].numpy() @ n - 0.1).max() < 1e-12
    # every face is oriented with the gradient: normal . n > 0
    f = o["faces"].numpy()
    nrm = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    assert np.all(nrm @ n > 0)


def test_sphere_mesh_is_closed_oriented_genus_zero():
    r = 0.55
    cubes, s, ci = grid_band(lambda P: np.linalg.norm(P, axis=1) - r, 16, 0.25)
    o = dmc_forward(torch.tensor(cubes), torch.tensor(s), torch.tensor(ci), 16)
    v, f = o["vertices"].numpy(), o["faces"].numpy()
    e = np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), axis=1)
    _, cnt = np.unique(e, axis=0, return_counts=True)
    assert np.all(cnt == 2)
    assert len(v) - len(cnt) + len(f) == 2
    vol = np.einsum("ij,ij->i", v[f[:, 0]], np.cross(v[f[:, 1]], v[f[:, 2]])).sum() / 6
    assert abs(vol - 4 / 3 * np.pi * r ** 3) < 0.07 * 4 / 3 * np.pi * r ** 3
    rad = np.linalg.norm(v, axis=1)
    assert np.abs(rad - r).max() < 2.0 / 16 * 0.2


def test_oracle_gradient_matches_central_differences():
    cubes, s, ci = grid_band(shapes_sdf, 8, 0.3)
    rng = np.random.default_rng(3)
    N, M = len(cubes), len(s)
    vc, cit = torch.tensor(cubes), torch.tensor(ci)
    args = [torch.tensor(s, requires_grad=True), torch.tensor(rng.uniform(-0.01, 0.01, (M, 3)), requires_grad=True),
            torch.tensor(rng.standard_normal((N, 8)), requires_grad=True),
            torch.tensor(rng.standard_normal((N, 12)), requires_grad=True),
            torch.tensor(rng.standard_normal((M, 2)), requires_grad=True)]
    o0 = dmc_forward(vc, args[0], cit, 8, deform=args[1], alpha=args[2], beta=args[3], v_features=args[4])
    W = {k: torch.tensor(rng.standard_normal(tuple(o0[k].shape))) for k in OUTS}

    def loss(*a):
        o = dmc_forward(vc, a[0], cit, 8, deform=a[1], alpha=a[2], beta=a[3], v_features=a[4])
        assert torch.equal(o["faces"], o0["faces"])
        return sum((W[k] * o[k]).sum() for k in OUTS)

    loss(*args).backward()
    for i, a in enumerate(args):
        # sdf steps only at corners with |sdf| >> h, so no corner changes sign
        pool = np.nonzero(np.abs(s) > 1e-3)[0] if i == 0 else np.arange(a.numel())
        idx = rng.choice(pool, 12, replace=False)
        h = 1e-6
        for j in idx:
            d = torch.zeros(a.numel(), dtype=torch.float64)
            d[j] = h
            d = d.view_as(a)
            ap = [x.detach() + (d if k == i else 0) for k, x in enumerate(args)]
            am = [x.detach() - (d if k == i else 0) for k, x in enumerate(args)]
            fd = (loss(*ap) - loss(*am)).item() / (2 * h)
            g = a.grad.view(-1)[j].item()
            assert abs(fd - g) <= 1e-5 * max(1.0, abs(g)), (i, j, fd, g)


QEF_FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dmc_qef_fixtures.npz")
QEF_OUTS = ["vertices", "L_dev", "p_alpha", "n_alpha", "v_features"]


def qef_inputs(q, case, variant, dt):
    """Keyword arguments of one QEF fixture variant (see
    tests/golden/make_dmc_qef_fixtures.py)."""
    kw = {}
    if variant == "deform":
        kw = dict(deform=torch.tensor(q[f"{case}/deform"], dtype=dt),
                  v_features=torch.tensor(q[f"{case}/v_features"], dtype=dt))
    if variant == "given":
        kw = dict(corner_gradients=torch.tensor(q[f"{case}/corner_gradients"], dtype=dt))
    return kw


@pytest.fixture(scope="module")
def qfx():
    return dict(np.load(QEF_FIX))


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("tag", ["f64", "f32"])
@pytest.mark.parametrize("variant", ["plain", "deform", "given"])
def test_oracle_qef_matches_reference_vectors(qfx, case, tag, variant):
    dt = torch.float64 if tag == "f64" else torch.float32
    key = f"{case}_{tag}_{variant}"
    o = dmc_forward(torch.tensor(qfx[f"{case}/voxel_coords"]), torch.tensor(qfx[f"{case}/sdf"], dtype=dt),
                    torch.tensor(qfx[f"{case}/cube_idx"]), int(qfx[f"{case}/res"]), dtype=dt, training=False,
                    mode="qef", **qef_inputs(qfx, case, variant, dt))
    assert np.array_equal(o["faces"].numpy(), qfx[f"{key}/out_faces"])
    atol = 1e-12 if tag == "f64" else 2e-6
    for k in QEF_OUTS:
        if f"{key}/out_{k}" not in qfx:
            assert o[k] is None
            continue
        np.testing.assert_allclose(o[k].numpy(), qfx[f"{key}/out_{k}"], rtol=0, atol=atol, err_msg=k)


def test_qef_single_crossing_closed_form():
    """One crossing p with unit normal n: the system is n.d = 0 plus
    sqrt(reg) d = sqrt(reg) t (t = centre - p), so d = t - (n.t) n / (1 + reg)."""
    from oracle.dmc import solve_qef

    rng = np.random.default_rng(4)
    nv, reg = 64, 0.01
    p = torch.tensor(rng.uniform(-0.1, 0.1, (nv, 3)))
    n = torch.tensor(rng.standard_normal((nv, 3)))
    n = n / n.norm(dim=1, keepdim=True)
    cen = torch.tensor(rng.uniform(-0.05, 0.05, (nv, 3)))
    big = torch.full((nv, 3), 10.0)
    vd = solve_qef(p, n, torch.arange(nv), nv, reg, cen, -big, big)
    t = cen - p
    want = p + t - (n * t).sum(1, keepdim=True) * n / (1 + reg)
    np.testing.assert_allclose(vd.numpy(), want.numpy(), rtol=0, atol=1e-14)
    # outside the cell (one cell of slack) -> the centroid, i.e. p itself
    vd = solve_qef(p, n, torch.arange(nv), nv, reg, cen, p - 1e-3, p + 1e-3)
    far = ((want - p).abs() > 2e-3).any(1)
    assert bool(far.any())
    np.testing.assert_array_equal(vd[far].numpy(), p[far].numpy())


def test_corner_gradients_of_a_linear_field_closed_form():
    """Axis-aligned edges of a linear field s = g.x all have slope g_a along
    their direction.  Edges x and z run from the lower to the upper corner,
    edges y (corners 2->0, 3->1, 7->5, 6->4) from the upper to the lower, and
    the reference adds +slope*dir at an edge's first corner and -slope*dir at
    its second, so corner gradient a = g_a * (first-end count - second-end
    count) / (edges along a + 1e-4): 0 inside, +-g_a / (1 + 1e-4) on faces."""
    from oracle.dmc import estimate_corner_gradients

    r = 5
    g = torch.tensor([0.7, -1.3, 0.4], dtype=torch.float64)
    cubes, s, ci = grid_band(lambda P: P @ g.numpy(), r, 10.0)  # every cube
    assert len(cubes) == r ** 3
    pos = np.zeros((len(s), 3))
    off = np.array([[(k >> 0) & 1, (k >> 1) & 1, (k >> 2) & 1] for k in range(8)])
    for k in range(8):
        pos[ci[:, k]] = cubes + off[k]
    world = torch.tensor(pos * (2.0 / r) - 1.0)
    cg = estimate_corner_gradients(world, torch.tensor(s), torch.tensor(ci), torch.float64).numpy()
    for a in range(3):
        lo, hi = pos[:, a] == 0, pos[:, a] == r
        n_a = np.where(lo | hi, 1.0, 2.0)
        first_minus_second = np.where(lo, 1.0, 0.0) - np.where(hi, 1.0, 0.0)
        if a == 1:
            first_minus_second = -first_minus_second
        want = g[a].item() * first_minus_second / (n_a + 1e-4)
        np.testing.assert_allclose(cg[:, a], want, rtol=0, atol=1e-12)

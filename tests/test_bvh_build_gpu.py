"""Structure of the device-built BVH (checked on the host before any query
test walks it)."""
import numpy as np
import pytest
import torch

from atom3d_amd import datasets
from atom3d_amd.kernels.bvh import BVHAccelerator

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


@pytest.mark.parametrize("nf", [0, 1, 3, 4, 5, 8, 9, 100, 20480])
def test_tree_structure_small(nf):
    V, F = datasets.noisy_icosphere(subdivisions=5, seed=0)
    F = F[:nf]
    print(f"[bvh build] {nf} faces", flush=True)
    b = BVHAccelerator(V.to(DEV), F.to(DEV))
    info = b.check_tree()
    assert info["n_tris"] == nf


def test_tree_structure_duplicates():
    # identical centroids (all-equal Morton keys) stress the Karras tie rule
    V = torch.tensor([[0, 0, 0], [1, 0, 0], [0, 1, 0]], dtype=torch.float32)
    F = torch.tensor([[0, 1, 2]] * 1000, dtype=torch.int32)
    b = BVHAccelerator(V.to(DEV), F.to(DEV))
    b.check_tree()


def test_tree_structure_sphere_union():
    V, F = datasets.sphere_union()
    print("[bvh build] sphere union 4.1M faces", flush=True)
    b = BVHAccelerator(V.to(DEV), F.to(DEV))
    info = b.check_tree()
    assert info["n_tris"] == 4096000
    print(info)


def _sah_cost(b, ci=1.2, ct=1.0, walk=0.03):
    """Cost of the emitted tree under the build's cost model (box measure =
    Steiner volume of the box grown by r = walk x scene diagonal, over r),
    from the blob on the host."""
    raw = b.blob.cpu().numpy()
    hdr = raw[:256].view(np.int32)
    n_nodes = int(hdr[2])
    nodes_off = int(raw[16:24].view(np.int64)[0])
    lo, hi = raw[32:44].view(np.float32).astype(np.float64), raw[44:56].view(np.float32).astype(np.float64)
    r = walk * np.linalg.norm(hi - lo)
    nodes = raw[nodes_off:nodes_off + 64 * n_nodes].view(np.float32).reshape(n_nodes, 16).astype(np.float64)
    refs = raw[nodes_off:nodes_off + 64 * n_nodes].view(np.int32).reshape(n_nodes, 16)[:, 12:14]

    def measure(bx):
        d = np.maximum(bx[:, 3:6] - bx[:, 0:3], 0.0)
        a = 2.0 * (d[:, 0] * d[:, 1] + d[:, 1] * d[:, 2] + d[:, 2] * d[:, 0])
        return d.prod(1) / r + a + np.pi * d.sum(1) * r + 4.0 / 3.0 * np.pi * r * r

    lb, rb = nodes[:, 0:6], nodes[:, 6:12]
    un = np.concatenate([np.minimum(lb[:, :3], rb[:, :3]), np.maximum(lb[:, 3:], rb[:, 3:])], 1)
    cost = ci * measure(un).sum()
    for side, bx in ((0, lb), (1, rb)):
        rr = refs[:, side]
        leaf = (rr < 0) & (bx[:, 0] <= bx[:, 3])
        cnt = ((~rr[leaf]) & 7) + 1
        cost += ct * (cnt * measure(bx[leaf])).sum()
    return cost


@pytest.mark.parametrize("mesh", ["noisy", "union"])
def test_treelet_refinement_valid_cheaper_same_results(mesh):
    if mesh == "noisy":
        V, F = datasets.noisy_icosphere(subdivisions=5, seed=0)
    else:
        V, F = datasets.sphere_union(n_spheres=20, subdivisions=4)
    print(f"[bvh treelet] {mesh} {F.shape[0]} faces", flush=True)
    b0 = BVHAccelerator(V.to(DEV), F.to(DEV), treelet_passes=0)
    b3 = BVHAccelerator(V.to(DEV), F.to(DEV), treelet_passes=3)
    i0, i3 = b0.check_tree(), b3.check_tree()
    assert i0["n_nodes"] == i3["n_nodes"]
    assert 1 <= i3["depth"] <= 96 and 1 <= i0["depth"] <= 96
    c0, c3 = _sah_cost(b0), _sah_cost(b3)
    print(f"tree cost plain {c0:.4g} refined {c3:.4g} ({c3 / c0:.3f}); depth {i0['depth']} -> {i3['depth']}")
    assert c3 < c0
    P = datasets.uniform_points(1 << 16, seed=4).to(DEV)
    for kw in (dict(), dict(flags=1), dict(flags=2)):
        r0, r3 = b0.udf(P, **kw), b3.udf(P, **kw)
        for x, y in zip(r0, r3):
            assert torch.equal(x, y), kw


def test_treelet_build_is_deterministic():
    V, F = datasets.sphere_union(n_spheres=20, subdivisions=4)
    parts = []
    for _ in range(2):
        raw = BVHAccelerator(V.to(DEV), F.to(DEV), treelet_passes=3).blob.cpu().numpy()
        hdr = raw[:256].view(np.int32)
        n_tris, n_nodes = int(hdr[1]), int(hdr[2])
        nodes_off, tris_off = (int(x) for x in raw[16:32].view(np.int64))
        # the node capacity past n_nodes is never written
        parts.append((raw[:256], raw[nodes_off:nodes_off + 64 * n_nodes], raw[tris_off:tris_off + 80 * n_tris]))
    for a, b in zip(*parts):
        np.testing.assert_array_equal(a, b)


def test_treelet_duplicates_and_small():
    V = torch.tensor([[0, 0, 0], [1, 0, 0], [0, 1, 0]], dtype=torch.float32)
    F = torch.tensor([[0, 1, 2]] * 1000, dtype=torch.int32)
    BVHAccelerator(V.to(DEV), F.to(DEV), treelet_passes=3).check_tree()
    V, F = datasets.noisy_icosphere(subdivisions=2, seed=0)
    for nf in (9, 17, 33, 100):
        BVHAccelerator(V.to(DEV), F[:nf].to(DEV), treelet_passes=3).check_tree()

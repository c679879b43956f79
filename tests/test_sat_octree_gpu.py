"""GPU parity of the SAT box queries and the octree voxelisation.

Bar: bit-exact hit masks, pair sets (sorted by (box, face)) and octree voxel
lists (same order) against the oracle's restatement of the reference: its
SAH tree with 4 triangles per leaf, a face kept when its leaf box overlaps
the query box and the SAT test passes (kernels/bvh_kernels.cu:555-610).
"""
import numpy as np
import pytest
import torch

import oracle
from atom3d_amd import MeshBVH, OctreeIndexer, datasets
from conftest import sphere_union_oracle

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _boxes(rng, n, lo=-1.0, hi=1.0, smin=1e-3, smax=0.3):
    c = rng.uniform(lo, hi, (n, 3))
    h = np.exp(rng.uniform(np.log(smin), np.log(smax), (n, 3)))
    return (c - h).astype(np.float32), (c + h).astype(np.float32)


def test_aabb_hit_and_pairs_bitexact():
    V, F = datasets.noisy_icosphere(subdivisions=4, seed=1)
    Vn, Fn = V.numpy(), F.numpy()
    rng = np.random.default_rng(0)
    qmin, qmax = _boxes(rng, 4096)
    print(f"[aabb parity] F={len(Fn)} K={len(qmin)}", flush=True)
    m = MeshBVH(V.to(DEV), F.to(DEV))
    r0 = m.intersect_aabb(torch.from_numpy(qmin).to(DEV), torch.from_numpy(qmax).to(DEV), mode=0)
    r1 = m.intersect_aabb(torch.from_numpy(qmin).to(DEV), torch.from_numpy(qmax).to(DEV), mode=1)
    sah = oracle.SahBvh(Vn, Fn)
    hit, pa, pf = sah.aabb(qmin, qmax)
    np.testing.assert_array_equal(r0.hit.cpu().numpy(), hit)
    np.testing.assert_array_equal(r1.hit.cpu().numpy(), hit)
    order = np.lexsort((pf, pa))
    np.testing.assert_array_equal(r1.aabb_ids.cpu().numpy(), pa[order])
    np.testing.assert_array_equal(r1.face_ids.cpu().numpy(), pf[order])
    assert r1.aabb_ids.dtype == torch.int32 and r1.face_ids.dtype == torch.int32
    # the leaf-box rule reaches beyond the triangles' own boxes (SAT slack)
    _, ca, cf = sah.aabb(qmin, qmax, canonical=True)
    extra = len(pa) - len(ca)
    print(f"[aabb] reference pairs {len(pa)}, of which {extra} fail the triangle's own box test")
    assert extra >= 0


def test_aabb_pairs_eps_band():
    """Boxes placed within the SAT slack (1e-6) of faces: the reference keeps
    a face whose own box misses the query box when its leaf's box overlaps."""
    V, F = datasets.noisy_icosphere(subdivisions=4, seed=3)
    Vn, Fn = V.numpy(), F.numpy()
    rng = np.random.default_rng(2)
    t = Vn[Fn]
    f = rng.integers(0, len(Fn), 6000)
    ax = rng.integers(0, 3, 6000)
    side = rng.integers(0, 2, 6000)
    gap = rng.uniform(0, 9e-7, 6000).astype(np.float32)
    tlo, thi = t[f].min(1), t[f].max(1)
    size = rng.uniform(1e-3, 3e-2, (6000, 3)).astype(np.float32)
    c = (tlo + thi) / 2
    lo, hi = (c - size).astype(np.float32), (c + size).astype(np.float32)
    r = np.arange(6000)
    # put the box just beyond the face's box along one axis
    lo[r[side == 0], ax[side == 0]] = thi[side == 0, ax[side == 0]] + gap[side == 0]
    hi[r[side == 0], ax[side == 0]] = lo[r[side == 0], ax[side == 0]] + size[side == 0, 0]
    hi[r[side == 1], ax[side == 1]] = tlo[side == 1, ax[side == 1]] - gap[side == 1]
    lo[r[side == 1], ax[side == 1]] = hi[r[side == 1], ax[side == 1]] - size[side == 1, 0]
    m = MeshBVH(V.to(DEV), F.to(DEV))
    got = m.intersect_aabb(torch.from_numpy(lo).to(DEV), torch.from_numpy(hi).to(DEV), mode=1)
    sah = oracle.SahBvh(Vn, Fn)
    hit, pa, pf = sah.leaf_brute_pairs(lo, hi)
    np.testing.assert_array_equal(got.hit.cpu().numpy(), hit)
    np.testing.assert_array_equal(got.aabb_ids.cpu().numpy(), pa)
    np.testing.assert_array_equal(got.face_ids.cpu().numpy(), pf)
    _, ca, _ = sah.aabb(lo, hi, canonical=True)
    print(f"[aabb eps band] {len(pa)} reference pairs, {len(pa) - len(ca)} beyond the faces' own boxes")
    assert len(pa) > len(ca)


def test_aabb_brute_force_small():
    V, F = datasets.noisy_icosphere(subdivisions=2, seed=2)
    rng = np.random.default_rng(1)
    qmin, qmax = _boxes(rng, 2000, smin=1e-4, smax=0.05)
    m = MeshBVH(V.to(DEV), F.to(DEV))
    r = m.intersect_aabb(torch.from_numpy(qmin).to(DEV), torch.from_numpy(qmax).to(DEV), mode=1)
    hit, pa, pf = oracle.SahBvh(V.numpy(), F.numpy()).leaf_brute_pairs(qmin, qmax)
    np.testing.assert_array_equal(r.hit.cpu().numpy(), hit)
    np.testing.assert_array_equal(r.aabb_ids.cpu().numpy(), pa)
    np.testing.assert_array_equal(r.face_ids.cpu().numpy(), pf)


def test_aabb_empty_inputs():
    V, F = datasets.noisy_icosphere(subdivisions=1)
    m = MeshBVH(V.to(DEV), F.to(DEV))
    e = torch.zeros(0, 3, device=DEV)
    assert m.intersect_aabb(e, e, mode=0).hit.shape == (0,)
    r = m.intersect_aabb(e, e, mode=1)
    assert r.hit.shape == (0,) and r.aabb_ids.shape == (0,)
    far = torch.full((4, 3), 5.0, device=DEV)
    r = m.intersect_aabb(far, far + 1, mode=1)
    assert not r.hit.any() and r.aabb_ids.numel() == 0


@pytest.mark.parametrize("min_level,max_level", [(2, 7), (4, 8), (0, 5), (6, 6)])
def test_octree_bitexact_icosphere(min_level, max_level):
    V, F = datasets.noisy_icosphere(subdivisions=5, seed=0)
    print(f"[octree parity] icosphere levels {min_level}->{max_level}", flush=True)
    m = MeshBVH(V.to(DEV), F.to(DEV))
    oc = OctreeIndexer(max_level=max_level, device=DEV)
    got = oc.octree_traverse(m, min_level=min_level).cpu().numpy()
    sah = oracle.SahBvh(V.numpy(), F.numpy())
    ref, tested, hits = sah.octree([-1, -1, -1, 1, 1, 1], min_level, max_level)
    assert got.dtype == np.int64
    np.testing.assert_array_equal(got, ref)


def test_octree_custom_bounds_and_empty():
    V, F = datasets.noisy_icosphere(subdivisions=3, seed=4)
    m = MeshBVH(V.to(DEV), F.to(DEV))
    b = torch.tensor([[-0.61, -0.55, -0.7], [0.52, 0.6, 0.49]])
    oc = OctreeIndexer(max_level=7, bounds=b.to(DEV))
    got = oc.octree_traverse(m, min_level=3).cpu().numpy()
    ref, _, _ = oracle.SahBvh(V.numpy(), F.numpy()).octree(b.numpy().reshape(-1), 3, 7)
    np.testing.assert_array_equal(got, ref)
    far = torch.tensor([[5.0, 5.0, 5.0], [6.0, 6.0, 6.0]], device=DEV)
    oc = OctreeIndexer(max_level=6, bounds=far)
    out = oc.octree_traverse(m, min_level=2)
    assert out.shape == (0, 3) and out.dtype == torch.int64


@pytest.mark.parametrize("max_level", [8, 10])
def test_octree_sphere_union(max_level):
    """The 4,096,000-face sphere union, levels 4 -> max_level (1024^3 at 10)."""
    print(f"[octree parity] sphere union 4.1M faces levels 4->{max_level}", flush=True)
    V, F, sah = sphere_union_oracle()
    m = MeshBVH(V.to(DEV), F.to(DEV))
    oc = OctreeIndexer(max_level=max_level, device=DEV)
    got = oc.octree_traverse(m, min_level=4).cpu().numpy()
    ref, _, hits = sah.octree([-1, -1, -1, 1, 1, 1], 4, max_level)
    print(f"[octree] {len(got)} voxels at level {max_level} (oracle {len(ref)}); per level {list(hits[4:max_level + 1])}")
    np.testing.assert_array_equal(got, ref)

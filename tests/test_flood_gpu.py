"""GPU parity of the flood fill against the oracle: dense labels, the sparse
pipeline's six outputs and sparse_flood_fill's dict, bit-exact and in order."""
import numpy as np
import pytest
import torch

import oracle
from oracle import flood as oflood
from atom3d_amd import MeshBVH, OctreeIndexer, datasets
from atom3d_amd.apps import get_dam_neighborhood_signs, sparse_flood_fill
from atom3d_amd.kernels.flood_fill import flood_fill_3d, flood_fill_3d_sparse

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


@pytest.mark.parametrize("shape,density,seed", [((37, 53, 150), 0.3, 0), ((16, 16, 128), 0.2, 1),
                                                ((100, 70, 260), 0.35, 2), ((5, 3, 7), 0.1, 3)])
@pytest.mark.parametrize("conn", [6, 18, 26])
def test_dense_labels_bitexact(shape, density, seed, conn):
    rng = np.random.default_rng(seed)
    d = density if conn == 6 else density + 0.15
    occ = rng.random(shape) < d
    free = np.argwhere(~occ)
    start = tuple(int(x) for x in free[rng.integers(len(free))])
    print(f"[flood dense] {shape} conn {conn} start {start}", flush=True)
    got = flood_fill_3d(torch.from_numpy(occ).to(DEV), start, conn).cpu().numpy()
    ref = oflood.flood_fill_dense(occ, start, conn)
    np.testing.assert_array_equal(got, ref)


def _icosphere_voxels(res, sub=3):
    V, F = datasets.noisy_icosphere(subdivisions=sub, seed=0)
    m = MeshBVH(V.to(DEV), F.to(DEV))
    oc = OctreeIndexer(max_level=int(np.log2(res)), device=DEV)
    return m, oc.octree_traverse(m, min_level=2)


@pytest.mark.parametrize("conn", [6, 18, 26])
def test_sparse_pipeline_bitexact(conn):
    m, vox = _icosphere_voxels(128)
    got = flood_fill_3d_sparse(vox, 128, (0, 0, 0), padding=2, connectivity=conn)
    ref = oflood.flood_fill_sparse(vox.cpu().numpy(), 128, (0, 0, 0), padding=2, connectivity=conn)
    names = ["water", "dry", "collision", "true_dam", "dam_mask", "extended"]
    for n, g, r in zip(names, got, ref):
        g = g.cpu().numpy()
        assert g.shape == r.shape, n
        np.testing.assert_array_equal(g, r, err_msg=n)
    assert got[0].dtype == torch.int64 and got[4].dtype == torch.bool


@pytest.mark.parametrize("conn,k_ring", [(26, 1), (6, 1), (26, 2), (18, 0)])
def test_sparse_flood_fill_dict_bitexact(conn, k_ring):
    m, vox = _icosphere_voxels(64)
    got = sparse_flood_fill(m, 64, (0, 0, 0), connectivity=conn, min_level=2, k_ring=k_ring)
    ref = oflood.sparse_flood_fill_from_voxels(vox.cpu().numpy(), 64, (0, 0, 0), conn, k_ring)
    for k in ref:
        np.testing.assert_array_equal(got[k].cpu().numpy(), ref[k], err_msg=k)
    assert got["extended_tide_ijk"].dtype == torch.int32
    # the cavity of a closed sphere stays dry: interior voxels exist
    assert len(got["dry_ijk"]) > 0 and len(got["water_ijk"]) > 0


def test_neighborhood_signs_function():
    rng = np.random.default_rng(4)
    dam = np.unique(rng.integers(0, 20, (200, 3)), axis=0)
    water = np.unique(rng.integers(0, 20, (300, 3)), axis=0)
    dry = np.unique(rng.integers(0, 20, (300, 3)), axis=0)
    nb, s = get_dam_neighborhood_signs(torch.from_numpy(dam).to(DEV), torch.from_numpy(water).to(DEV),
                                       torch.from_numpy(dry).to(DEV), 20, k_ring=2)
    rnb, rs = oflood.neighborhood_signs(dam, water, dry, 20, 2)
    np.testing.assert_array_equal(nb.cpu().numpy(), rnb)
    np.testing.assert_array_equal(s.cpu().numpy(), rs)


def test_occupied_source_and_empty():
    dam = torch.tensor([[3, 3, 3], [3, 3, 4]], device=DEV)
    out = flood_fill_3d_sparse(dam, 16, source=(3, 3, 3))
    assert all(len(o) == 0 for o in out) and out[0].dtype == torch.int32
    out = flood_fill_3d_sparse(torch.zeros(0, 3, dtype=torch.int64, device=DEV), 16)
    assert all(len(o) == 0 for o in out)


def test_sphere_union_res256_bitexact():
    V, F = datasets.sphere_union()
    m = MeshBVH(V.to(DEV), F.to(DEV))
    print("[flood] sphere union res 256", flush=True)
    got = sparse_flood_fill(m, 256, (0, 0, 0), connectivity=26, min_level=4)
    vox = OctreeIndexer(max_level=8, device=DEV).octree_traverse(m, min_level=4)
    ref = oflood.sparse_flood_fill_from_voxels(vox.cpu().numpy(), 256, (0, 0, 0), 26, 1)
    for k in ref:
        np.testing.assert_array_equal(got[k].cpu().numpy(), ref[k], err_msg=k)

"""Voxel-to-dual-graph glue on the GPU (kernels/cells.py through the C ABI)
against the numpy oracle (oracle/cells.py): unique cells, unique corner
vertices with the cube -> vertex table, world positions, the 6-neighbour
table with the reference's unchecked flat keys, wet face / edge / corner
masks, construct_surface_dual_graph on a real flood fill and
voxel_unique_vertices on the demo's octree pipeline.  Integer outputs
bit-exact, world positions bit-exact (same two float32 roundings)."""
import numpy as np
import pytest
import torch

from atom3d_amd import MeshBVH, OctreeIndexer, datasets
from atom3d_amd.apps import (construct_surface_dual_graph, get_dam_wet_corners, get_dam_wet_edges,
                             get_dam_wet_faces, sparse_flood_fill)
from atom3d_amd.kernels import cells as kc
from oracle import cells as oc

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


@pytest.mark.parametrize("method", [1, 2])
@pytest.mark.parametrize("R,n", [(16, 3000), (1024, 200000), (1700, 50000), (100, 5)])
def test_unique_cells_vertices_neighbors_match_oracle(R, n, method):
    """Both methods (1 radix sort, 2 bitmaps); R = 1700 takes the 64-bit key
    path of the sort ((R+1)^3 > 2^32); R = 16 and 100 have partial 32-bit
    words in the bitmap rows."""
    print(f"cells R={R} n={n} method={method}", flush=True)
    rng = np.random.default_rng(R)
    base = rng.integers(0, R, (n // 4, 3))
    cells = np.concatenate([base, base[rng.integers(0, len(base), n - len(base))]])  # duplicates
    cells = np.concatenate([cells, np.clip(cells[: n // 8] + rng.integers(-1, 2, (n // 8, 3)), 0, R - 1)])
    t = torch.tensor(cells, dtype=torch.int32, device=DEV)
    u = kc.unique_cells(t, R, method=method)
    ref_u = oc.unique_cells(cells)
    assert np.array_equal(u.cpu().numpy(), ref_u)
    lo, cell = [-1.0, -0.5, -2.0], [float(np.float32(2.0) / np.float32(R))] * 3
    cube, vkey, vijk, world = kc.cell_vertices(u, R, lo=lo, cell=cell, want_keys=True, want_ijk=True, method=method)
    ref_v, ref_cube = oc.cell_vertices(ref_u)
    assert np.array_equal(vijk.cpu().numpy(), ref_v)
    assert np.array_equal(cube.cpu().numpy(), ref_cube)
    P = R + 1
    assert np.array_equal(vkey.cpu().numpy(), (ref_v[:, 0] * P + ref_v[:, 1]) * P + ref_v[:, 2])
    assert np.array_equal(world.cpu().numpy(), oc.world(ref_v, lo, cell))
    adj = kc.cell_neighbors(u, R, method=method)
    assert np.array_equal(adj.cpu().numpy(), oc.cell_neighbors(ref_u, R))
    # unsorted cells with duplicates straight into cell_vertices
    cube2, _, vijk2, _ = kc.cell_vertices(t, R, want_ijk=True, method=method)
    ref_v2, ref_cube2 = oc.cell_vertices(cells)
    assert np.array_equal(vijk2.cpu().numpy(), ref_v2) and np.array_equal(cube2.cpu().numpy(), ref_cube2)


@pytest.mark.parametrize("R", [24, 1024])
def test_wet_masks_match_oracle(R):
    print(f"wet R={R}", flush=True)
    rng = np.random.default_rng(5)
    n = 4000 if R == 24 else 60000
    centre = rng.integers(0, R, (n, 3))
    dam = np.unique(centre, axis=0)
    coll = np.unique(np.clip(centre + rng.integers(-1, 2, (n, 3)), 0, R - 1), axis=0)
    coll = coll[rng.random(len(coll)) < 0.5]
    # include the grid faces
    dam = np.unique(np.concatenate([dam, [[0, 0, 0], [R - 1, R - 1, R - 1], [0, R - 1, 3]]]), axis=0)
    d, c = torch.tensor(dam, device=DEV, dtype=torch.int32), torch.tensor(coll, device=DEV, dtype=torch.int64)
    if R <= 64:
        for fn, ref in ((get_dam_wet_faces, oc.wet_faces), (get_dam_wet_edges, oc.wet_edges),
                        (get_dam_wet_corners, oc.wet_corners)):
            got = fn(d, c, R)
            assert got.dtype == torch.bool
            assert np.array_equal(got.cpu().numpy(), ref(dam, coll, R)), fn.__name__
    else:  # dense oracle mask too large: check against a sorted-key lookup
        keys = np.sort((coll[:, 0] * R + coll[:, 1]) * R + coll[:, 2])
        for fn, table in ((get_dam_wet_faces, oc.NEIGHBOR_OFFSETS_6.numpy()[:, None, :]),
                          (get_dam_wet_edges, oc.EDGE_NEIGHBOR_OFFSETS.numpy()),
                          (get_dam_wet_corners, oc.CORNER_NEIGHBOR_OFFSETS.numpy())):
            nb = dam[:, None, None, :] + table[None].astype(np.int64)
            ok = ((nb >= 0) & (nb < R)).all(-1)
            k = (nb[..., 0] * R + nb[..., 1]) * R + nb[..., 2]
            pos = np.clip(np.searchsorted(keys, k), 0, len(keys) - 1)
            want = ((keys[pos] == k) & ok).any(-1)
            assert np.array_equal(fn(d, c, R).cpu().numpy(), want), fn.__name__
    assert get_dam_wet_corners(d[:0], c, R).shape == (0, 8)
    assert not get_dam_wet_faces(d, c[:0], R).any()
    with pytest.raises(IndexError):
        get_dam_wet_faces(d, torch.tensor([[0, 0, R]], device=DEV), R)


def test_dual_graph_of_a_flood_fill_matches_oracle():
    print("dual graph of a 128^3 flood fill", flush=True)
    V, F = datasets.noisy_icosphere(subdivisions=4, seed=0)
    m = MeshBVH(V.to(DEV), F.to(DEV))
    fl = sparse_flood_fill(m, 128, connectivity=26, min_level=2, verbose=False)
    oct_ = OctreeIndexer(max_level=7, device=DEV)
    dam, coll, dry = fl["dam_ijk"], fl["tide_ijk"], fl["dry_adjacent_ijk"]
    assert len(dam) > 1000 and len(dry) > 100
    w, vijk, cube, adj, act = construct_surface_dual_graph(dam, coll, dry, oct_)
    rw, rv, rcube, radj, ract = oc.construct_surface_dual_graph(dam.cpu().numpy(), coll.cpu().numpy(),
                                                                dry.cpu().numpy(), 128, [-1.0] * 3, [1.0] * 3)
    assert act.dtype == torch.promote_types(dam.dtype, coll.dtype) and vijk.dtype == torch.int64
    assert cube.dtype == torch.int64 and adj.dtype == torch.int64 and w.dtype == torch.float32
    assert np.array_equal(act.cpu().numpy(), ract)
    assert np.array_equal(vijk.cpu().numpy(), rv)
    assert np.array_equal(cube.cpu().numpy(), rcube)
    assert np.array_equal(adj.cpu().numpy(), radj)
    assert np.array_equal(w.cpu().numpy(), rw)
    # dry_adjacent None and a single list
    w1, v1, c1, a1, act1 = construct_surface_dual_graph(dam, coll[:0], None, oct_)
    assert np.array_equal(act1.cpu().numpy(), oc.unique_cells(dam.cpu().numpy()))
    with pytest.raises(RuntimeError, match="empty"):
        construct_surface_dual_graph(dam[:0], coll[:0], None, oct_)


def test_voxel_unique_vertices_demo_pipeline():
    """demo.ipynb cells 1-2: icosphere(subdivisions=1), octree to 512^3 over
    the mesh bounds, voxel_unique_vertices of the candidates, UDF at the
    vertices with unit-norm gradient."""
    print("voxel_unique_vertices demo pipeline", flush=True)
    V, F = datasets.icosphere(subdivisions=1)
    bvh = MeshBVH(torch.tensor(V, dtype=torch.float32, device=DEV), torch.tensor(F, dtype=torch.int32, device=DEV))
    gi = OctreeIndexer(max_level=9, bounds=bvh.get_bounds(), device=DEV)
    ijk = gi.octree_traverse(bvh)
    idx = gi.ijk_to_cube(ijk)
    uv, coords, mapping = gi.voxel_unique_vertices(idx)
    P = 513
    cand = (ijk.cpu().numpy()[:, None, :] + oc._CORNERS[None]).reshape(-1, 3)
    keys = (cand[:, 0] * P + cand[:, 1]) * P + cand[:, 2]
    ref_u, ref_inv = np.unique(keys, return_inverse=True)
    assert uv.dtype == torch.int64 and mapping.dtype == torch.int64 and coords.dtype == torch.float32
    assert np.array_equal(uv.cpu().numpy(), ref_u)
    assert np.array_equal(mapping.cpu().numpy(), ref_inv.reshape(-1, 8))
    vijk = np.stack([ref_u // (P * P), (ref_u // P) % P, ref_u % P], 1)
    b = gi.bounds.cpu().numpy()
    assert np.array_equal(coords.cpu().numpy(), oc.world(vijk, b[0], gi.cell_size.cpu().numpy()))
    # [N,3] input gives the same
    uv2, coords2, mapping2 = gi.voxel_unique_vertices(ijk)
    assert torch.equal(uv2, uv) and torch.equal(mapping2, mapping) and torch.equal(coords2, coords)
    x = coords.clone().requires_grad_()
    d = bvh.udf(x, return_grad=True)
    g = torch.autograd.grad(d.sum(), x)[0]
    far = d.detach() > 1e-4
    assert torch.allclose(g.norm(dim=1)[far], torch.ones_like(d[far]), atol=1e-4)
    for method in (1, 2):
        with pytest.raises(RuntimeError, match="outside"):
            kc.cell_vertices(torch.tensor([[0, 0, 600]], device=DEV, dtype=torch.int32), 512, method=method)

"""Voxelizer on the GPU (implicit candidate block, first-hit SAT cells,
pair CSR) against the oracle (oracle/voxelizer.py): voxel coordinates and
their order, face lists, starts and counts bit-exact, for the candidate and
all-cells strategies, default and mesh-fitted bounds."""
import numpy as np
import pytest
import torch

from atom3d_amd import CubeGrid, MeshBVH, datasets
from atom3d_amd.apps import Voxelizer
from oracle import voxelizer as ov

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _case(name):
    if name == "ico1_fit":
        v, f = datasets.icosphere(1, 1.0)
        return v.astype(np.float32), f.astype(np.int32), True
    V, F = datasets.noisy_icosphere(subdivisions=3, seed=4)
    return V.numpy().astype(np.float32), F.numpy().astype(np.int32), False


@pytest.mark.parametrize("name,res", [("ico1_fit", 40), ("noisy3", 48), ("noisy3", 17)])
def test_voxelize_with_faces_matches_oracle(name, res):
    print(f"voxelizer {name} res={res}", flush=True)
    V, F, fit = _case(name)
    m = MeshBVH(torch.tensor(V, device=DEV), torch.tensor(F, device=DEV))
    grid = CubeGrid(res, bounds=m.get_bounds() if fit else None, device=DEV)
    vz = Voxelizer(m, grid)
    b0, cell = grid.bounds[0].cpu().numpy(), grid.cell_size.cpu().numpy()
    vf = vz.voxelize_with_faces()
    ref = ov.voxelize_with_faces(V, F, b0, cell, res)
    assert vf.voxel_coords.dtype == torch.int64 and vf.face_indices.dtype == torch.int32
    assert vf.face_start.dtype == torch.int32 and vf.face_count.dtype == torch.int32
    for got, want in zip((vf.voxel_coords, vf.face_indices, vf.face_start, vf.face_count), ref):
        assert np.array_equal(got.cpu().numpy(), want)
    assert torch.equal(vf.get_faces_for_voxel(3).cpu(), torch.tensor(ref[1][ref[2][3]:ref[2][3] + ref[3][3]]))
    surf = vz.voxelize_surface()
    assert torch.equal(surf, vf.voxel_coords)
    all_ = vz.voxelize_surface(strategy="all")
    ref_all = ov.voxelize_with_faces(V, F, b0, cell, res, strategy="all")[0]
    assert np.array_equal(all_.cpu().numpy(), ref_all)
    pm = vz.voxelize_with_polygons()
    P = int(vf.face_indices.shape[0])
    assert pm.polygons.shape == (P, 8, 3) and pm.polygon_counts.shape == (P,)
    assert torch.equal(pm.polygons[:, :3], m.vertices[m.faces[vf.face_indices.long()].long()])
    assert (pm.polygons[:, 3:] == 0).all() and (pm.polygon_counts == 3).all()
    assert torch.equal(pm.voxel_ids.long(), torch.repeat_interleave(torch.arange(len(vf.face_count), device=DEV),
                                                                    vf.face_count.long()))
    assert torch.equal(pm.get_polygon(0), pm.polygons[0, :3])


def test_candidate_cells_equal_reference_formula():
    V, F, _ = _case("noisy3")
    m = MeshBVH(torch.tensor(V, device=DEV), torch.tensor(F, device=DEV))
    grid = CubeGrid(33, device=DEV)
    lo, hi = m.get_face_aabb()
    cand = grid.generate_candidate_cells_from_aabb(lo, hi)
    o, d = ov.candidate_block(V, F, grid.bounds[0].cpu().numpy(), grid.cell_size.cpu().numpy(), 33)
    assert np.array_equal(cand.cpu().numpy(), ov.block_cells(o, d))
    assert grid.generate_all_cells().shape == (33 ** 3, 3)

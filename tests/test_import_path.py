"""The reference's import paths resolve to this build's modules (no GPU
needed: the HIP library is loaded on the first call, not on import)."""
import importlib
import sys


def test_reference_import_paths():
    import atom3d
    import atom3d_amd
    from atom3d import CubeGrid, MeshBVH, OctreeIndexer  # atom3d/__init__.py:11-21
    from atom3d.apps import Voxelizer, construct_surface_dual_graph, get_dam_wet_faces, sparse_flood_fill
    from atom3d.grid import OctreeIndexer as O2
    from atom3d.mesh_extractor import SparseDiffDMC, SparseDMC

    assert MeshBVH is atom3d_amd.MeshBVH and O2 is OctreeIndexer and CubeGrid is atom3d_amd.CubeGrid
    assert SparseDiffDMC is atom3d_amd.mesh_extractor.SparseDiffDMC and SparseDMC is not None
    assert sparse_flood_fill is atom3d_amd.apps.sparse_flood_fill
    assert Voxelizer and construct_surface_dual_graph and get_dam_wet_faces
    for name in ("atom3d.apps.sparse_flood_fill", "atom3d.mesh_extractor.sparse_diffdmc",
                 "atom3d.grid.octree_indexer", "atom3d.core.mesh_bvh", "atom3d.kernels.bvh"):
        m = importlib.import_module(name)
        assert m is sys.modules["atom3d_amd" + name[len("atom3d"):]]
    assert set(atom3d.__all__) >= {"MeshBVH", "OctreeIndexer", "CubeGrid", "ClosestPointResult",
                                   "AABBIntersectResult"}

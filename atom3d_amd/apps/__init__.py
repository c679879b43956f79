"""Apps on the hot path (same import path as atom3d.apps)."""
from .sparse_flood_fill import (construct_surface_dual_graph, get_dam_neighborhood_signs, get_dam_wet_corners,
                                get_dam_wet_edges, get_dam_wet_faces, sparse_flood_fill,
                                sparse_flood_fill_from_voxels)
from .voxelizer import Voxelizer

__all__ = ["sparse_flood_fill", "get_dam_neighborhood_signs", "get_dam_wet_faces", "get_dam_wet_edges",
           "get_dam_wet_corners", "construct_surface_dual_graph", "sparse_flood_fill_from_voxels", "Voxelizer"]

"""Grid module exports (same names as atom3d.grid)."""
from .cube_grid import CubeGrid
from .octree_indexer import OctreeIndexer

__all__ = ["OctreeIndexer", "CubeGrid"]

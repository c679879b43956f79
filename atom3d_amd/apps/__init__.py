"""Apps on the hot path (same import path as atom3d.apps)."""
from .sparse_flood_fill import get_dam_neighborhood_signs, sparse_flood_fill

__all__ = ["sparse_flood_fill", "get_dam_neighborhood_signs"]

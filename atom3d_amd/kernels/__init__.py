"""Native kernels of atom3d_amd (gfx950 HIP library bound with ctypes)."""
from ._lib import LIB_PATH, SIGNATURES, lib, version  # noqa: F401
from .bvh import BVHAccelerator, bvh_available  # noqa: F401

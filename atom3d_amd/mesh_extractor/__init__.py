"""Differentiable mesh extraction (Atom3D atom3d/mesh_extractor/__init__.py):
SparseDiffDMC / SparseDMC, sparse dual marching cubes with HIP kernels for
the topology, the weighted-average dual vertices and their adjoint."""
from .sparse_diffdmc import SparseDiffDMC, SparseDiffDMCOutput, SparseDMC

__all__ = ["SparseDiffDMC", "SparseDMC", "SparseDiffDMCOutput"]

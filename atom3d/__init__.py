"""The reference's import path for this build.

Code written against Atom3D (``from atom3d import MeshBVH``,
``from atom3d.apps import sparse_flood_fill``,
``from atom3d.mesh_extractor import SparseDiffDMC``,
``from atom3d.grid import OctreeIndexer``; atom3d/__init__.py:11-21,
apps/__init__.py:1-39, mesh_extractor/__init__.py:1-12) runs unchanged:
every module of the atom3d_amd package is registered under the same dotted
name with ``atom3d`` in front, as the SAME module object, so classes,
functions and module state are shared between the two names.
"""
import importlib
import pkgutil
import sys

import atom3d_amd as _impl
from atom3d_amd import *  # noqa: F401,F403  (the reference's top-level names)

__all__ = list(_impl.__all__)
__version__ = _impl.__version__


def _register():
    for info in pkgutil.walk_packages(_impl.__path__, prefix="atom3d_amd."):
        if info.name.rsplit(".", 1)[-1] in ("build",) or ".csrc" in info.name:
            continue
        mod = importlib.import_module(info.name)
        alias = "atom3d" + info.name[len("atom3d_amd"):]
        sys.modules.setdefault(alias, mod)
        parent, _, leaf = alias.rpartition(".")
        setattr(sys.modules[parent], leaf, mod)


_register()

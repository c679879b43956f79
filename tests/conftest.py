import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


def _has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if _has_gpu():
        return
    skip = pytest.mark.skip(reason="no GPU visible")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)


_CACHE = {}


def sphere_union_oracle():
    """(V, F, oracle.SahBvh) of the 4,096,000-face sphere union, built once
    per session (the oracle's single-threaded SAH build takes ~30 s)."""
    if "c2" not in _CACHE:
        import oracle
        from atom3d_amd import datasets

        V, F = datasets.sphere_union()
        _CACHE["c2"] = (V, F, oracle.SahBvh(V.numpy(), F.numpy()))
    return _CACHE["c2"]

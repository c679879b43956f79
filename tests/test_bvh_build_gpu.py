"""Structure of the device-built BVH (checked on the host before any query
test walks it)."""
import numpy as np
import pytest
import torch

from atom3d_amd import datasets
from atom3d_amd.kernels.bvh import BVHAccelerator

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


@pytest.mark.parametrize("nf", [0, 1, 3, 4, 5, 8, 9, 100, 20480])
def test_tree_structure_small(nf):
    V, F = datasets.noisy_icosphere(subdivisions=5, seed=0)
    F = F[:nf]
    print(f"[bvh build] {nf} faces", flush=True)
    b = BVHAccelerator(V.to(DEV), F.to(DEV))
    info = b.check_tree()
    assert info["n_tris"] == nf


def test_tree_structure_duplicates():
    # identical centroids (all-equal Morton keys) stress the Karras tie rule
    V = torch.tensor([[0, 0, 0], [1, 0, 0], [0, 1, 0]], dtype=torch.float32)
    F = torch.tensor([[0, 1, 2]] * 1000, dtype=torch.int32)
    b = BVHAccelerator(V.to(DEV), F.to(DEV))
    b.check_tree()


def test_tree_structure_sphere_union():
    V, F = datasets.sphere_union()
    print("[bvh build] sphere union 4.1M faces", flush=True)
    b = BVHAccelerator(V.to(DEV), F.to(DEV))
    info = b.check_tree()
    assert info["n_tris"] == 4096000
    print(info)

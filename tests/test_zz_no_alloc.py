"""a3d_query_order_noalloc (workspace in the library's fixed device scratch)
against a3d_query_order (caller workspace): the same permutation bit for
bit, on two calls in a row, the second with fewer queries (the scratch
keeps the first call's contents); then the ordered closest-point walk on
it equals the walk on a3d_query_order's permutation."""
import ctypes

import pytest
import torch

from atom3d_amd import datasets
from atom3d_amd.kernels import _lib
from atom3d_amd.kernels.bvh import BVHAccelerator

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _order_with_ws(P):
    K = P.shape[0]
    wsb = ctypes.c_int64()
    _lib.call("a3d_query_order_ws_bytes", K, ctypes.byref(wsb))
    ws = torch.empty(max(wsb.value, 1), dtype=torch.uint8, device=DEV)
    perm = torch.empty(K, dtype=torch.int32, device=DEV)
    _lib.call("a3d_query_order", _lib.ptr(P), K, _lib.ptr(perm), _lib.ptr(ws), wsb.value, _lib.stream(P.device))
    return perm


def _order_noalloc(P):
    perm = torch.full((P.shape[0],), -1, dtype=torch.int32, device=DEV)
    _lib.call("a3d_query_order_noalloc", _lib.ptr(P), P.shape[0], _lib.ptr(perm), _lib.stream(P.device))
    return perm


def test_noalloc_order_equals_workspace_order():
    V, F = datasets.noisy_icosphere(subdivisions=4, seed=2)
    b = BVHAccelerator(V.to(DEV), F.to(DEV))
    for K, seed in ((1 << 21, 1), (100003, 2)):  # second call smaller: the scratch holds the first's data
        print(f"[noalloc] K={K}", flush=True)
        P = datasets.uniform_points(K, seed=seed, lo=-0.7, hi=0.9).to(DEV)
        ref = _order_with_ws(P)
        got = _order_noalloc(P)
        torch.cuda.synchronize()
        assert torch.equal(ref, got)
        assert torch.equal(torch.sort(got.long()).values, torch.arange(K, device=DEV))
        outs = []
        for perm in (ref, got):
            o = [torch.empty(K, device=DEV), torch.empty(K, dtype=torch.int32, device=DEV),
                 torch.empty(K, 3, device=DEV), torch.empty(K, 3, device=DEV)]
            _lib.call("a3d_bvh_udf", _lib.ptr(b.blob), _lib.ptr(P), K, _lib.ptr(perm), *[_lib.ptr(t) for t in o], 0,
                      None, _lib.stream(P.device))
            outs.append(o)
        for x, y in zip(*outs):
            assert torch.equal(x, y)


def test_noalloc_limits():
    P = torch.zeros(8, 3, device=DEV)
    _lib.call("a3d_query_order_noalloc", _lib.ptr(P), 0, None, _lib.stream(P.device))  # K = 0: nothing to do
    with pytest.raises(RuntimeError, match="scratch"):
        _lib.call("a3d_query_order_noalloc", _lib.ptr(P), (1 << 21) + 1, _lib.ptr(P), _lib.stream(P.device))

"""Oracle (test infrastructure only): the hash-grid + MLP UDF surrogate
restated with CPU torch ops.

There is no reference counterpart; this restates the design the HIP kernels
implement (atom3d_amd/csrc/surrogate.hip): 16 levels with resolutions
floor(16 * (2048/16)^(l/15)), dense indexing where (N+1)^3 <= 2^19 entries,
else the spatial hash (x * 1) ^ (y * 2654435761) ^ (z * 805459861) mod 2^19,
trilinear interpolation of 2 features per level; MLP 32 -> 64 -> 64 -> 1
with ReLU, bf16 operands (round to nearest even) and fp32 sums; loss
mean((y - t)^2).  `bf16=False` turns the roundings off, which is how the
manual backward is pinned against torch autograd.
"""
from __future__ import annotations

import math

import torch

L, LOG2T = 16, 19
T = 1 << LOG2T
N_MLP = 64 * 32 + 64 * 64 + 64 + 1
PRIMES = (1, 2654435761, 805459861)


def level_layout():
    b = (2048.0 / 16.0) ** (1.0 / (L - 1))
    N = [int(math.floor(16.0 * b ** l + 1e-9)) for l in range(L)]
    dense = [(n + 1) ** 3 <= T for n in N]
    return N, dense


def corner_index(lv: int, ix, iy, iz):
    N, dense = level_layout()
    if dense[lv]:
        r = N[lv] + 1
        return ix + r * (iy + r * iz)
    h = (ix.to(torch.int64) * PRIMES[0]) ^ (iy.to(torch.int64) * PRIMES[1]) ^ (iz.to(torch.int64) * PRIMES[2])
    return h & (T - 1)


def _cells(P, lo, hi):
    """Per level: corner ids [B, 8] and trilinear weights [B, 8] (float32)."""
    N, _ = level_layout()
    P = P.float()
    inv = (1.0 / (torch.tensor(hi, dtype=torch.float32) - torch.tensor(lo, dtype=torch.float32))).float()
    x = ((P - torch.tensor(lo, dtype=torch.float32)) * inv).clamp(0.0, 1.0)
    out = []
    for lv in range(L):
        pos = x * float(N[lv])
        i0 = pos.to(torch.int64).clamp(0, N[lv] - 1)
        w = pos - i0.float()
        ids, ws = [], []
        for k in range(8):
            b = [(k >> d) & 1 for d in range(3)]
            ids.append(corner_index(lv, i0[:, 0] + b[0], i0[:, 1] + b[1], i0[:, 2] + b[2]))
            wx = w[:, 0] if b[0] else 1.0 - w[:, 0]
            wy = w[:, 1] if b[1] else 1.0 - w[:, 1]
            wz = w[:, 2] if b[2] else 1.0 - w[:, 2]
            ws.append((wx * wy) * wz)
        out.append((torch.stack(ids, 1), torch.stack(ws, 1)))
    return out


def encode(P, table, lo, hi, cells=None):
    """Features [B, 32]: feature 2*level + c."""
    cells = _cells(P, lo, hi) if cells is None else cells
    feats = []
    for lv, (ids, ws) in enumerate(cells):
        v = table[lv][ids]  # [B, 8, 2]
        w = ws.unsqueeze(-1).to(v.dtype)
        f = w[:, 0] * v[:, 0]
        for k in range(1, 8):  # corner order, as the kernels sum
            f = f + w[:, k] * v[:, k]
        feats.append(f)
    return torch.cat(feats, 1)


def _r(x, on):
    return x.bfloat16().to(x.dtype) if on else x


def unpack(mlp):
    W1 = mlp[:2048].reshape(64, 32)
    W2 = mlp[2048:6144].reshape(64, 64)
    return W1, W2, mlp[6144:6208], mlp[6208]


def forward(P, table, mlp, lo, hi, bf16=True):
    """(y [B], intermediates) of the network."""
    W1, W2, w3, b3 = unpack(mlp)
    X = _r(encode(P, table, lo, hi), bf16)
    H1 = X @ _r(W1, bf16).T
    H1b = _r(torch.relu(H1), bf16)
    H2 = H1b @ _r(W2, bf16).T
    y = torch.relu(H2) @ w3 + b3
    return y, dict(X=X, H1=H1, H1b=H1b, H2=H2)


def grads(P, target, table, mlp, lo, hi, bf16=True):
    """Manual backward of mean((y - t)^2) with the kernels' rounding points:
    (loss, mlp gradient [6209], table gradient [L, T, 2])."""
    W1, W2, w3, b3 = unpack(mlp)
    cells = _cells(P, lo, hi)
    X = _r(encode(P, table, lo, hi, cells), bf16)
    H1 = X @ _r(W1, bf16).T
    H1b = _r(torch.relu(H1), bf16)
    H2 = H1b @ _r(W2, bf16).T
    y = torch.relu(H2) @ w3 + b3
    B = P.shape[0]
    diff = y - target
    loss = (diff * diff).sum() / B
    dy = 2.0 * diff / B
    on2 = (H2 > 0).to(H2.dtype)
    gw3 = (dy[:, None] * torch.relu(H2)).sum(0)
    gb3 = dy.sum()
    dH2 = _r(dy[:, None] * w3[None, :] * on2, bf16)
    dH1 = _r((dH2 @ _r(W2, bf16)) * (H1 > 0).to(H1.dtype), bf16)
    dX = dH1 @ _r(W1, bf16)
    gW2 = dH2.T @ H1b
    gW1 = dH1.T @ X
    gmlp = torch.cat([gW1.reshape(-1), gW2.reshape(-1), gw3, gb3.reshape(1)])
    gtab = torch.zeros_like(table)
    for lv, (ids, ws) in enumerate(cells):
        g = ws.unsqueeze(-1).to(dX.dtype) * dX[:, None, 2 * lv:2 * lv + 2]  # [B, 8, 2]
        gtab[lv].index_add_(0, ids.reshape(-1), g.reshape(-1, 2))
    return loss, gmlp, gtab


def init_like_device(seed: int = 0):
    """Small deterministic parameters for CPU-only checks (not the device's RNG)."""
    g = torch.Generator().manual_seed(seed)
    table = (torch.rand(L, T, 2, generator=g) * 2 - 1) * 1e-1
    mlp = torch.cat([(torch.rand(2048, generator=g) * 2 - 1) * math.sqrt(6 / 96),
                     (torch.rand(4096, generator=g) * 2 - 1) * math.sqrt(6 / 128),
                     (torch.rand(64, generator=g) * 2 - 1) * math.sqrt(6 / 65), torch.zeros(1)])
    return table, mlp


__all__ = ["level_layout", "corner_index", "encode", "forward", "grads", "unpack", "init_like_device", "N_MLP"]

"""The surrogate oracle (oracle/surrogate.py) pinned on the CPU: its manual
backward equals torch autograd of its forward (float64, roundings off), its
trilinear hash-grid encoding reproduces a linear field exactly on the dense
levels, the spatial hash matches 32-bit modular arithmetic done with Python
integers, and the level layout is 16..2048."""
import numpy as np
import torch

from oracle import surrogate as S


def test_level_layout():
    N, dense = S.level_layout()
    assert N[0] == 16 and N[-1] == 2048 and len(N) == 16
    assert all(b > a for a, b in zip(N, N[1:]))
    assert dense == [(n + 1) ** 3 <= 2 ** 19 for n in N] and dense[0] and not dense[-1]


def test_hash_matches_integer_arithmetic():
    rng = np.random.default_rng(0)
    ijk = rng.integers(0, 2049, (1000, 3))
    got = S.corner_index(15, *(torch.tensor(ijk[:, d]) for d in range(3))).numpy()
    want = [((int(i) * 1) ^ ((int(j) * 2654435761) % 2 ** 32) ^ ((int(k) * 805459861) % 2 ** 32)) % 2 ** 19
            for i, j, k in ijk]
    assert np.array_equal(got, want)


def test_encoding_reproduces_linear_field_on_dense_levels():
    N, dense = S.level_layout()
    table = torch.zeros(S.L, S.T, 2, dtype=torch.float64)
    a = torch.tensor([0.3, -0.7, 1.1], dtype=torch.float64)
    for lv in range(S.L):
        if not dense[lv]:
            continue
        r = N[lv] + 1
        idx = torch.arange(r ** 3)
        ix, iy, iz = idx % r, (idx // r) % r, idx // (r * r)
        g = torch.stack([ix, iy, iz], 1).double() / N[lv]  # grid point in unit coordinates
        table[lv, :r ** 3, 0] = g @ a
        table[lv, :r ** 3, 1] = 1.0
    P = torch.rand(500, 3, dtype=torch.float64) * 1.8 - 0.9
    f = S.encode(P, table, [-1.0] * 3, [1.0] * 3)
    x = ((P.float() + 1) * 0.5).double()
    for lv in range(S.L):
        if dense[lv]:
            assert torch.allclose(f[:, 2 * lv], x @ a, atol=1e-5)
            assert torch.allclose(f[:, 2 * lv + 1], torch.ones(500, dtype=torch.float64), atol=1e-6)


def test_manual_backward_equals_autograd():
    table, mlp = S.init_like_device(1)
    table, mlp = table.double(), mlp.double()
    g = torch.Generator().manual_seed(2)
    P = torch.rand(300, 3, generator=g, dtype=torch.float64) * 2 - 1
    t = torch.rand(300, generator=g, dtype=torch.float64)
    loss, gm, gt = S.grads(P, t, table, mlp, [-1.0] * 3, [1.0] * 3, bf16=False)
    tab = table.clone().requires_grad_()
    m = mlp.clone().requires_grad_()
    y, _ = S.forward(P, tab, m, [-1.0] * 3, [1.0] * 3, bf16=False)
    l2 = ((y - t) ** 2).mean()
    l2.backward()
    assert torch.allclose(loss, l2.detach())
    assert torch.allclose(gm, m.grad, rtol=1e-10, atol=1e-12)
    assert torch.allclose(gt, tab.grad, rtol=1e-10, atol=1e-12)
    assert gt.abs().sum() > 0 and gm.abs().sum() > 0

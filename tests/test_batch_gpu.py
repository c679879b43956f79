"""Batched uneven closest-point queries (one persistent launch over many
meshes) against the per-mesh walk, bit for bit; graph replays against the
eager run; every autotuner launch shape gives the same results."""
import pytest
import torch

from atom3d_amd import datasets
from atom3d_amd.batch import BatchedUDF

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


@pytest.fixture(scope="module")
def small_batch():
    meshes = datasets.uneven_batch(n_meshes=12, seed=5, tri_max=30000, q_min=1, q_max=1 << 13)
    meshes[3] = (meshes[3][0], meshes[3][1], 0)  # a mesh without queries
    P, counts = datasets.batch_queries(meshes, seed=5, device=DEV)
    bu = BatchedUDF([(V, F) for V, F, _ in meshes], device=DEV)
    return meshes, P, counts, bu


def _per_mesh(bu, P, counts):
    outs = []
    o = 0
    for b, c in zip(bu.bvhs, counts):
        outs.append(b.udf(P[o:o + c]))
        o += c
    return [torch.cat([x[k] for x in outs]) for k in range(4)]


def _same(a, b):
    return all(torch.equal(x, y) for x, y in zip(a, b))


def test_batched_equals_per_mesh(small_batch):
    meshes, P, counts, bu = small_batch
    print(f"[batch] {len(meshes)} meshes, {sum(counts)} queries, faces {bu.n_faces}", flush=True)
    plan = bu.plan(P, counts)
    got = [t.clone() for t in bu.run(plan, {"waves_per_block": 4, "waves_per_cu": 16, "order": "cost"})]
    torch.cuda.synchronize()
    ref = _per_mesh(bu, P, counts)
    assert _same(got[:3], ref[:3])
    has = ref[1] >= 0
    assert torch.equal(got[3][has], ref[3][has])


def test_short_bounds_are_walked_again(small_batch):
    """Bound records whose grid distances are halved and that carry no
    closest-face bound fall short of many queries' distances: the bounded walk
    reports those as unfound and the redo pass walks them again, so the
    results still equal the per-mesh walk bit for bit."""
    import numpy as np

    from atom3d_amd.batch import _BOUND

    meshes, P, counts, bu = small_batch
    assert bu.bounds is not None
    rec = bu.bounds.cpu().numpy().view(_BOUND).copy()
    halves = []
    for m, b in enumerate(bu.bvhs):
        h = (b.seed_grid(bu.BOUND_GRID_RES)[0] * 0.5).contiguous()
        halves.append(h)
        rec[m]["G"] = h.data_ptr()
        rec[m]["GF"] = 0
    saved = bu.bounds
    bu.bounds = torch.from_numpy(rec.view(np.uint8).copy()).to(bu.device)
    try:
        plan = bu.plan(P, counts)
        got = [t.clone() for t in bu.run(plan, {"waves_per_block": 4, "waves_per_cu": 16, "order": "cost"})]
        torch.cuda.synchronize()
    finally:
        bu.bounds = saved
    ref = _per_mesh(bu, P, counts)
    assert _same(got[:3], ref[:3])
    has = ref[1] >= 0
    assert torch.equal(got[3][has], ref[3][has])


def test_graph_replays_match_eager(small_batch):
    meshes, P, counts, bu = small_batch
    plan = bu.plan(P, counts)
    cfg = {"waves_per_block": 2, "waves_per_cu": 8, "order": "mesh"}
    eager = [t.clone() for t in bu.run(plan, cfg)]
    g = bu.capture(plan, cfg)
    for r in range(3):
        for t in plan.out:
            t.zero_()
        plan.perm.zero_()
        torch.cuda.synchronize()
        g.replay()
        g.stream.synchronize()
        print(f"[batch graph] replay {r}", flush=True)
        assert _same(plan.out[:3], eager[:3])
        _ = torch.ones(1000, device=DEV).sum().item()  # eager work between replays
    g.close()


def test_autotune_shapes_agree(small_batch):
    meshes, P, counts, bu = small_batch
    plan = bu.plan(P, counts)
    best, table = bu.autotune(plan, reps=1, cache=False)
    assert best in [c for c, _ in table] and all(ms > 0 for _, ms in table)
    ref = [t.clone() for t in bu.run(plan, table[0][0])]
    for cfg, _ in table[1:]:
        out = bu.run(plan, cfg)
        assert _same(out[:3], ref[:3]), cfg


def test_autotune_choice_persists(small_batch, tmp_path, monkeypatch):
    """The second autotune of the same shape bucket reads the JSON cache and
    launches nothing."""
    meshes, P, counts, bu = small_batch
    path = str(tmp_path / "autotune.json")
    plan = bu.plan(P, counts)
    best, table = bu.autotune(plan, reps=1, cache=path)
    key = bu.tune_key(plan)
    assert key.startswith("gfx950/") and "/k_udf_batch/" in key
    calls = []
    monkeypatch.setattr(bu, "_launch", lambda *a, **k: calls.append(a))
    best2, table2 = bu.autotune(plan, reps=1, cache=path)
    assert (best2, table2) == (best, table) and calls == []


def test_plan_rejects_bad_counts(small_batch):
    meshes, P, counts, bu = small_batch
    with pytest.raises(ValueError):
        bu.plan(P, counts[:-1])

"""Host-side configuration (no GPU): the autotuner's JSON cache (round trip,
replacement, damaged and foreign entries) and the roctx ranges around C-ABI
calls (loaded only when ATOM3D_ROCTX is set)."""
import json
import os
import subprocess
import sys

from atom3d_amd.batch import autotune_cache_path, default_tune_space, load_tuned, store_tuned

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_tune_cache_round_trip(tmp_path):
    path = str(tmp_path / "sub" / "autotune.json")
    assert load_tuned(path, "k") is None  # no file yet
    space = default_tune_space()
    table = [(c, float(i + 1)) for i, c in enumerate(space)]
    store_tuned(path, "gfx950/cu256/k_udf_batch/K2^26", space[3], table)
    store_tuned(path, "other", space[0], table[:2])
    cfg, tab = load_tuned(path, "gfx950/cu256/k_udf_batch/K2^26")
    assert cfg == space[3] and tab == table
    store_tuned(path, "other", space[5], table[:1])  # replaces, keeps the first key
    assert load_tuned(path, "other")[0] == space[5]
    assert load_tuned(path, "gfx950/cu256/k_udf_batch/K2^26")[0] == space[3]
    assert [p for p in os.listdir(tmp_path / "sub")] == ["autotune.json"]  # no temporary files left


def test_tune_cache_rejects_damaged_entries(tmp_path):
    path = str(tmp_path / "autotune.json")
    with open(path, "w") as fh:
        fh.write("{not json")
    assert load_tuned(path, "k") is None
    store_tuned(path, "k", default_tune_space()[0], [])  # a damaged file is replaced
    assert load_tuned(path, "k")[0] == default_tune_space()[0]
    bad = {"k1": {"config": {"waves_per_block": 3, "waves_per_cu": 8, "order": "cost"}, "table": []},
           "k2": {"config": {"waves_per_block": "2", "waves_per_cu": 8, "order": "cost"}, "table": []},
           "k3": {"table": []}, "k4": [1, 2]}
    with open(path, "w") as fh:
        json.dump(bad, fh)
    for k in bad:
        assert load_tuned(path, k) is None, k


def test_tune_cache_path_from_environment(monkeypatch, tmp_path):
    monkeypatch.setenv("ATOM3D_AUTOTUNE_CACHE", str(tmp_path / "x.json"))
    assert autotune_cache_path() == str(tmp_path / "x.json")
    monkeypatch.delenv("ATOM3D_AUTOTUNE_CACHE")
    assert autotune_cache_path().endswith(os.path.join(".cache", "atom3d_amd", "autotune.json"))


def _run(code, **env):
    e = dict(os.environ, **env)
    return subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=e, capture_output=True, text=True, timeout=300)


def test_roctx_ranges_only_when_asked():
    code = ("from atom3d_amd.kernels import _lib\n"
            "tr = _lib._roctx()\n"
            "print(bool(tr))\n"
            "if tr:\n"
            "    assert tr[0](b'outer') == 0 and tr[0](b'inner') == 1\n"
            "    assert tr[1]() == 1 and tr[1]() == 0\n"
            "    with _lib.trace_range('block'):\n"
            "        assert tr[0](b'nested') == 1\n"
            "        tr[1]()\n")
    on = _run(code, ATOM3D_ROCTX="1")
    assert on.returncode == 0, on.stderr
    assert on.stdout.strip() == "True"
    off = _run(code, ATOM3D_ROCTX="0")
    assert off.returncode == 0, off.stderr
    assert off.stdout.strip() == "False"

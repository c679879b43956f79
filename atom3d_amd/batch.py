"""Batched closest-point queries over many meshes of uneven size in one
persistent launch, captured in a HIP graph, with a launch-shape autotuner.
No counterpart in the reference (one mesh per MeshBVH call).

    bu = BatchedUDF([(V0, F0), (V1, F1), ...])
    plan = bu.plan(P_concat, counts)          # Morton order + tile table
    cfg, table = bu.autotune(plan)            # waves/block, waves/CU, tile order
    dist, face, closest, uvw = bu.run(plan, cfg)
    g = bu.capture(plan, cfg); g.replay()     # one graph launch per batch

The autotuner's choice is kept in a JSON file keyed by (GPU architecture and
CU count, kernel, shape bucket), so a process that meets a shape bucket
already tuned on this machine launches at once.
"""
from __future__ import annotations

import ctypes
import json
import math
import os
import tempfile
from dataclasses import dataclass, field

import numpy as np
import torch

from .kernels import _lib
from .kernels.bvh import BVHAccelerator

TILE = 64


@dataclass
class BatchPlan:
    P: torch.Tensor
    counts: list
    q_off: torch.Tensor      # device int64 [n+1]
    perm: torch.Tensor       # device uint32-as-int32 [K]
    ws: torch.Tensor
    tiles: dict = field(default_factory=dict)  # order -> device int64 [n_tiles, 2]
    out: tuple = ()
    counter: torch.Tensor = None

    @property
    def K(self) -> int:
        return int(self.P.shape[0])


class BatchGraph:
    """An instantiated HIP graph of one batch (order + persistent walk)."""

    def __init__(self, handle: int, stream: torch.cuda.Stream):
        self.handle = handle
        self.stream = stream

    def replay(self) -> None:
        _lib.call("a3d_graph_launch", self.handle, self.stream.cuda_stream)

    def close(self) -> None:
        if self.handle:
            _lib.call("a3d_graph_destroy", self.handle)
            self.handle = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# one a3d_grid_bound record (64 bytes) per mesh
_BOUND = np.dtype([("G", "<u8"), ("GF", "<u8"), ("V", "<u8"), ("F", "<u8"), ("R", "<i4"), ("lo", "<f4", 3),
                   ("h", "<f4"), ("margin", "<f4"), ("pad", "<i4", 2)])
assert _BOUND.itemsize == 64


class BatchedUDF:
    BOUND_GRID_RES = 63  # per-mesh distance-bound grid (64^3 points), built with the trees

    def __init__(self, meshes, device="cuda", leaf: int = 8, bounds: bool = True):
        self.device = torch.device(device)
        self.bvhs = [BVHAccelerator(torch.as_tensor(V).to(self.device), torch.as_tensor(F).to(self.device),
                                    n_primitives_per_leaf=leaf) for V, F in meshes]
        self.n_faces = [b.num_faces for b in self.bvhs]
        self.blobs = torch.tensor([b.blob.data_ptr() for b in self.bvhs], dtype=torch.int64, device=self.device)
        props = torch.cuda.get_device_properties(self.device)
        self.n_cu = int(props.multi_processor_count)
        self.bounds = self._bounds() if bounds else None

    def _bounds(self):
        """Device array of one grid-bound record per mesh (see
        BVHAccelerator.seed_grid), or None when a mesh has none (empty or
        non-finite)."""
        rec = np.zeros(len(self.bvhs), _BOUND)
        for m, b in enumerate(self.bvhs):
            sg = b.seed_grid(self.BOUND_GRID_RES) if b.num_faces > 0 else None
            if sg is None:
                return None
            g, gf, lo, h = sg
            rec[m] = (g.data_ptr(), gf.data_ptr(), b.vertices.data_ptr(), b.faces.data_ptr(), self.BOUND_GRID_RES,
                      lo, h, h / 1024, (0, 0))
        return torch.from_numpy(rec.view(np.uint8).copy()).to(self.device)

    # ------------------------------------------------------------------ plan
    def plan(self, P: torch.Tensor, counts) -> BatchPlan:
        P = P.to(self.device, torch.float32).contiguous()
        counts = [int(c) for c in counts]
        if len(counts) != len(self.bvhs) or sum(counts) != P.shape[0]:
            raise ValueError("counts must give one query count per mesh, summing to len(P)")
        K = int(P.shape[0])
        off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        q_off = torch.tensor(off, device=self.device)
        perm = torch.empty(max(K, 1), dtype=torch.int32, device=self.device)
        wsb = ctypes.c_int64()
        _lib.call("a3d_batch_order_ws_bytes", K, ctypes.byref(wsb))
        ws = torch.empty(max(wsb.value, 1), dtype=torch.uint8, device=self.device)
        out = (torch.empty(K, device=self.device), torch.empty(K, dtype=torch.int32, device=self.device),
               torch.empty(K, 3, device=self.device), torch.empty(K, 3, device=self.device))
        plan = BatchPlan(P=P, counts=counts, q_off=q_off, perm=perm, ws=ws, out=out,
                         counter=torch.zeros(1, dtype=torch.int32, device=self.device))
        self._order(plan, torch.cuda.current_stream(self.device))
        return plan

    def _order(self, plan: BatchPlan, stream) -> None:
        with torch.cuda.device(self.device):
            _lib.call("a3d_batch_order", _lib.ptr(self.blobs), len(self.bvhs), _lib.ptr(plan.P), plan.K,
                      _lib.ptr(plan.q_off), _lib.ptr(plan.perm), _lib.ptr(plan.ws), plan.ws.numel(),
                      stream.cuda_stream)

    def tiles(self, plan: BatchPlan, order: str = "cost") -> torch.Tensor:
        """Tile table: each mesh's sorted range cut into packets of <= 64;
        order 'mesh' keeps mesh order, 'cost' puts the meshes with the most
        expensive queries (largest trees) first."""
        if order in plan.tiles:
            return plan.tiles[order]
        off = np.concatenate([[0], np.cumsum(plan.counts)]).astype(np.int64)
        meshes = range(len(plan.counts))
        if order == "cost":
            meshes = sorted(meshes, key=lambda m: -math.log2(max(self.n_faces[m], 2)))
        elif order != "mesh":
            raise ValueError(f"unknown tile order {order!r}")
        rows = []
        for m in meshes:
            s = np.arange(off[m], off[m + 1], TILE, dtype=np.int64)
            c = np.minimum(TILE, off[m + 1] - s)
            rows.append(np.stack([(s << 8) | c, np.full_like(s, m)], 1))
        t = np.concatenate(rows) if rows else np.zeros((0, 2), np.int64)
        plan.tiles[order] = torch.tensor(t, device=self.device)
        return plan.tiles[order]

    # ------------------------------------------------------------------ run
    def _launch(self, plan: BatchPlan, cfg: dict, stream) -> None:
        t = self.tiles(plan, cfg.get("order", "cost"))
        wpb = int(cfg.get("waves_per_block", 4))
        blocks = max(1, self.n_cu * int(cfg.get("waves_per_cu", 16)) // wpb)
        d, f, c, w = plan.out
        _lib.call("a3d_batch_udf", _lib.ptr(self.blobs), _lib.ptr(plan.P), _lib.ptr(plan.perm), _lib.ptr(t),
                  int(t.shape[0]), _lib.ptr(plan.counter), wpb, blocks, _lib.ptr(d), _lib.ptr(f), _lib.ptr(c),
                  _lib.ptr(w), _lib.ptr(self.bounds), stream.cuda_stream)

    def run(self, plan: BatchPlan, cfg: dict | None = None, reorder: bool = False):
        """(dist, face, closest, uvw) of all queries, in input order."""
        cfg = cfg or {}
        s = torch.cuda.current_stream(self.device)
        with torch.cuda.device(self.device):
            if reorder:
                self._order(plan, s)
            self._launch(plan, cfg, s)
        return plan.out

    def capture(self, plan: BatchPlan, cfg: dict | None = None) -> BatchGraph:
        """One graph per batch: Morton ordering + persistent walk."""
        cfg = cfg or {}
        self.tiles(plan, cfg.get("order", "cost"))  # tables exist before capture
        stream = torch.cuda.Stream(self.device)
        stream.wait_stream(torch.cuda.current_stream(self.device))
        h = ctypes.c_int64()
        with torch.cuda.device(self.device):
            _lib.call("a3d_graph_begin", stream.cuda_stream)
            try:
                self._order(plan, stream)
                self._launch(plan, cfg, stream)
            finally:
                _lib.call("a3d_graph_end", stream.cuda_stream, ctypes.byref(h))
        return BatchGraph(h.value, stream)

    # ------------------------------------------------------------------ autotune
    def tune_key(self, plan: BatchPlan) -> str:
        """Cache key of a plan's launch-shape choice: the GPU (architecture
        and CU count), the kernel, and the shape bucket (log2 of the query
        count, the face count and the mesh count, rounded; whether per-mesh
        distance bounds are used)."""
        props = torch.cuda.get_device_properties(self.device)
        arch = str(getattr(props, "gcnArchName", "") or props.name).split(":")[0]

        def lg(x):
            return int(round(math.log2(max(int(x), 1))))

        return (f"{arch}/cu{self.n_cu}/k_udf_batch/K2^{lg(plan.K)}/F2^{lg(sum(self.n_faces))}"
                f"/M2^{lg(len(self.bvhs))}/bounds{int(self.bounds is not None)}")

    def autotune(self, plan: BatchPlan, configs=None, reps: int = 3,
                 cache: bool | str | None = True) -> tuple[dict, list]:
        """Time each launch shape on the plan and return the fastest with the
        table of (config, ms).  With the default search space and ``cache``
        true (or a file path), the choice is kept in a JSON file keyed by
        ``tune_key`` (default file: $ATOM3D_AUTOTUNE_CACHE, else
        ~/.cache/atom3d_amd/autotune.json) and a later call with the same key
        returns it without timing, with the stored table."""
        path = None
        if configs is None and cache:
            path = cache if isinstance(cache, str) else autotune_cache_path()
            hit = load_tuned(path, self.tune_key(plan))
            if hit is not None:
                return hit
        best, table = self._time_configs(plan, configs, reps)
        if path is not None:
            store_tuned(path, self.tune_key(plan), best, table)
        return best, table

    def _time_configs(self, plan: BatchPlan, configs, reps: int) -> tuple[dict, list]:
        if configs is None:
            configs = default_tune_space()
        s = torch.cuda.current_stream(self.device)
        table = []
        for cfg in configs:
            with torch.cuda.device(self.device):
                self._launch(plan, cfg, s)  # warm-up (and tile table)
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(s)
                for _ in range(reps):
                    self._launch(plan, cfg, s)
                e1.record(s)
            torch.cuda.synchronize(self.device)
            table.append((cfg, e0.elapsed_time(e1) / reps))
        best = min(table, key=lambda r: r[1])[0]
        return best, table


# ---------------------------------------------------------------------- autotune cache
_TUNE_KEYS = {"waves_per_block": int, "waves_per_cu": int, "order": str}


def default_tune_space() -> list:
    """The launch shapes the autotuner times: waves per block x resident waves
    per CU x tile order."""
    return [{"waves_per_block": wpb, "waves_per_cu": wpc, "order": order}
            for order in ("cost", "mesh") for wpb in (2, 4, 8, 16) for wpc in (8, 16, 32) if wpb <= wpc]


def autotune_cache_path() -> str:
    return os.environ.get("ATOM3D_AUTOTUNE_CACHE") or os.path.join(
        os.path.expanduser("~"), ".cache", "atom3d_amd", "autotune.json")


def _valid_cfg(cfg) -> bool:
    return (isinstance(cfg, dict) and set(cfg) == set(_TUNE_KEYS)
            and all(isinstance(cfg[k], t) for k, t in _TUNE_KEYS.items())
            and cfg in default_tune_space())


def load_tuned(path: str, key: str):
    """(config, table) stored under ``key``, or None when the file is absent,
    unreadable, or holds no valid entry for the key (a damaged cache only
    costs a re-tune)."""
    try:
        with open(path) as fh:
            ent = json.load(fh).get(key)
        cfg, table = ent["config"], [(c, float(ms)) for c, ms in ent["table"]]
    except (OSError, ValueError, KeyError, TypeError, AttributeError):
        return None
    return (cfg, table) if _valid_cfg(cfg) else None


def store_tuned(path: str, key: str, cfg: dict, table: list) -> None:
    """Add or replace one entry; the file is rewritten whole through a
    temporary file and an atomic rename, so concurrent writers never leave a
    torn file (the last one wins)."""
    try:
        with open(path) as fh:
            data = json.load(fh)
        if not isinstance(data, dict):
            data = {}
    except (OSError, ValueError):
        data = {}
    data[key] = {"config": cfg, "table": [[c, ms] for c, ms in table]}
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    fd, tmp = tempfile.mkstemp(prefix=".autotune.", dir=os.path.dirname(os.path.abspath(path)))
    try:
        with os.fdopen(fd, "w") as fh:
            json.dump(data, fh, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except BaseException:
        if os.path.exists(tmp):
            os.unlink(tmp)
        raise


__all__ = ["BatchedUDF", "BatchPlan", "BatchGraph", "TILE", "default_tune_space", "autotune_cache_path",
           "load_tuned", "store_tuned"]

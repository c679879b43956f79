"""Batched closest-point queries over many meshes of uneven size in one
persistent launch, captured in a HIP graph, with a launch-shape autotuner.
No counterpart in the reference (one mesh per MeshBVH call).

    bu = BatchedUDF([(V0, F0), (V1, F1), ...])
    plan = bu.plan(P_concat, counts)          # Morton order + tile table
    cfg, table = bu.autotune(plan)            # waves/block, waves/CU, tile order
    dist, face, closest, uvw = bu.run(plan, cfg)
    g = bu.capture(plan, cfg); g.replay()     # one graph launch per batch
"""
from __future__ import annotations

import ctypes
import math
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
    BOUND_GRID_RES = 31  # per-mesh distance-bound grid (32^3 points), built with the trees

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
    def autotune(self, plan: BatchPlan, configs=None, reps: int = 3) -> tuple[dict, list]:
        """Time each launch shape on the plan and return the fastest with the
        table of (config, ms)."""
        if configs is None:
            configs = [{"waves_per_block": wpb, "waves_per_cu": wpc, "order": order}
                       for order in ("cost", "mesh") for wpb in (2, 4, 8, 16) for wpc in (8, 16, 32)
                       if wpb <= wpc]
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


__all__ = ["BatchedUDF", "BatchPlan", "BatchGraph", "TILE"]

"""Learned UDF surrogate with exact BVH fallback (no reference counterpart).

A multiresolution hash grid + MLP (atom3d_amd/csrc/surrogate.hip) is trained
on the GPU by its own forward/backward kernels on UDF samples the BVH
labels, then answers distance queries: every query predicted inside a
safety ``band`` of the surface is re-answered exactly by the BVH walk, the
rest take the network's value.  The exact answers are the BVH's, bit for
bit; the network's error on the others is measured, not bounded, and is
reported by ``evaluate``.

    sg = UDFSurrogate(mesh)                     # MeshBVH or BVHAccelerator
    sg.fit(n_samples=1 << 24, steps=2000)
    d, face, closest, uvw, n_exact = sg.udf(points, band=0.01)
"""
from __future__ import annotations

import ctypes
import hashlib
import json
import time
from typing import Optional

import torch

from .kernels import _lib

# parameter blob layout (mirrors the kernels)
L, LOG2T = 16, 19
T = 1 << LOG2T
_OFF_TAB16 = 512
_OFF_W1B = _OFF_TAB16 + L * T * 4
_OFF_W2B = _OFF_W1B + 64 * 32 * 2
_OFF_W3F = _OFF_W2B + 64 * 64 * 2
N_MLP = 6209
GRAD_STRIDE = 6272


def _f3(v):
    return (ctypes.c_float * 3)(*[float(x) for x in v])


class UDFSurrogate:
    def __init__(self, mesh, bounds: Optional[torch.Tensor] = None, seed: int = 0):
        bvh = mesh._bvh if hasattr(mesh, "_bvh") else mesh
        self.bvh = bvh
        self.vertices = (mesh.vertices if hasattr(mesh, "vertices") else bvh.vertices).float().contiguous()
        self.faces = (mesh.faces if hasattr(mesh, "faces") else bvh.faces).int().contiguous()
        self.device = self.vertices.device
        _lib.require_cuda(self.vertices, "UDFSurrogate mesh")
        if bounds is None:
            bounds = torch.tensor([[-1.0, -1.0, -1.0], [1.0, 1.0, 1.0]])
        b = torch.as_tensor(bounds, dtype=torch.float32).reshape(2, 3).cpu()
        self.lo, self.hi = b[0].tolist(), b[1].tolist()
        pb, sb = ctypes.c_int64(), ctypes.c_int64()
        _lib.call("a3d_sg_sizes", ctypes.byref(pb), ctypes.byref(sb))
        self.params = torch.empty(pb.value, dtype=torch.uint8, device=self.device)
        self.state = torch.empty(sb.value, dtype=torch.uint8, device=self.device)
        self.seed = int(seed)
        self.step = 0
        with torch.cuda.device(self.device):
            _lib.call("a3d_sg_init", _lib.ptr(self.params), _lib.ptr(self.state), _f3(self.lo), _f3(self.hi),
                      self.seed, self._s())

    def _s(self):
        return _lib.stream(self.device)

    # ---------------------------------------------------------------- views
    def table_fp32(self) -> torch.Tensor:
        """The fp32 master hash table [16, 2^19, 2] (a view into the state)."""
        return self.state[: L * T * 8].view(torch.float32).view(L, T, 2)

    def mlp_fp32(self) -> torch.Tensor:
        """The fp32 master MLP vector [6209]: W1 [64,32], W2 [64,64], w3 [64], b3."""
        off = ((L * T * 8 + 255) // 256 * 256) * 4
        return self.state[off: off + N_MLP * 4].view(torch.float32)

    def table_fp16(self) -> torch.Tensor:
        return self.params[_OFF_TAB16:_OFF_W1B].view(torch.float16).view(L, T, 2)

    # ---------------------------------------------------------------- training
    def sample(self, n: int, near_frac: float = 0.5, radius: float = 0.02, seed: int = 1) -> torch.Tensor:
        P = torch.empty(n, 3, dtype=torch.float32, device=self.device)
        with torch.cuda.device(self.device):
            _lib.call("a3d_sg_sample", _lib.ptr(self.vertices), _lib.ptr(self.faces), int(self.faces.shape[0]), n,
                      float(near_frac), float(radius), _f3(self.lo), _f3(self.hi), int(seed), _lib.ptr(P), self._s())
        return P

    def label(self, P: torch.Tensor) -> torch.Tensor:
        return self.bvh.udf(P)[0]

    def grad(self, P: torch.Tensor, target: torch.Tensor) -> torch.Tensor:
        """Accumulate the gradient of one batch; returns the MLP gradient
        vector (6272 floats, the batch loss at index 6209)."""
        g = torch.empty(GRAD_STRIDE, dtype=torch.float32, device=self.device)
        with torch.cuda.device(self.device):
            _lib.call("a3d_sg_grad", _lib.ptr(self.params), _lib.ptr(self.state), _lib.ptr(P.contiguous()),
                      _lib.ptr(target.contiguous()), int(P.shape[0]), _lib.ptr(g), self._s())
        return g

    def table_grad(self) -> torch.Tensor:
        out = torch.empty(L, T, 2, dtype=torch.float32, device=self.device)
        with torch.cuda.device(self.device):
            _lib.call("a3d_sg_table_grad", _lib.ptr(self.state), _lib.ptr(out), self._s())
        return out

    def apply(self, lr: float, beta1: float = 0.9, beta2: float = 0.99, eps: float = 1e-15) -> None:
        self.step += 1
        with torch.cuda.device(self.device):
            _lib.call("a3d_sg_apply", _lib.ptr(self.params), _lib.ptr(self.state), float(lr), float(beta1),
                      float(beta2), float(eps), self.step, self._s())

    def fit(self, n_samples: int = 1 << 24, steps: int = 2000, batch: int = 1 << 18, lr: float = 1e-2,
            near_frac: float = 0.5, radius: float = 0.02, seed: int = 1, log_every: int = 0) -> dict:
        """Sample and label n_samples points, then `steps` Adam steps on
        batches drawn in order (reshuffled each epoch)."""
        t0 = time.perf_counter()
        P = self.sample(n_samples, near_frac, radius, seed)
        y = self.label(P)
        torch.cuda.synchronize(self.device)
        t1 = time.perf_counter()
        gen = torch.Generator(device=self.device).manual_seed(seed)
        order = torch.randperm(n_samples, device=self.device, generator=gen)
        losses = []
        pos = 0
        for it in range(steps):
            if pos + batch > n_samples:
                order = torch.randperm(n_samples, device=self.device, generator=gen)
                pos = 0
            idx = order[pos:pos + batch]
            pos += batch
            lr_t = lr * (0.1 ** (it / max(steps, 1)))
            g = self.grad(P[idx], y[idx])
            self.apply(lr_t)
            if log_every and (it % log_every == 0 or it == steps - 1):
                losses.append((it, float(g[N_MLP].item())))
        torch.cuda.synchronize(self.device)
        return {"sample_s": t1 - t0, "train_s": time.perf_counter() - t1, "steps": steps, "batch": batch,
                "losses": losses}

    # ---------------------------------------------------------------- queries
    def predict(self, P: torch.Tensor, perm: Optional[torch.Tensor] = None) -> torch.Tensor:
        P = P.contiguous().float()
        out = torch.empty(P.shape[0], dtype=torch.float32, device=self.device)
        with torch.cuda.device(self.device):
            _lib.call("a3d_sg_infer", _lib.ptr(self.params), _lib.ptr(P), int(P.shape[0]), _lib.ptr(perm),
                      _lib.ptr(out), self._s())
        return out

    def udf(self, P: torch.Tensor, band: float = 0.01, order: bool = True):
        """(dist, face, closest, uvw, n_exact): queries predicted within
        `band` come from the BVH (exact), the rest from the network (face -1)."""
        P = P.contiguous().float()
        K = int(P.shape[0])
        dist = torch.empty(K, dtype=torch.float32, device=self.device)
        face = torch.empty(K, dtype=torch.int32, device=self.device)
        closest = torch.empty(K, 3, dtype=torch.float32, device=self.device)
        uvw = torch.empty(K, 3, dtype=torch.float32, device=self.device)
        wsb = ctypes.c_int64()
        _lib.call("a3d_sg_route_ws_bytes", K, ctypes.byref(wsb))
        ws = torch.empty(max(wsb.value, 1), dtype=torch.uint8, device=self.device)
        perm = self.bvh.query_order(P) if order and K >= 4096 else None
        n = ctypes.c_int64()
        with torch.cuda.device(self.device):
            _lib.call("a3d_sg_udf", _lib.ptr(self.params), _lib.ptr(self.bvh.blob), _lib.ptr(P), K, _lib.ptr(perm),
                      float(band), _lib.ptr(dist), _lib.ptr(face), _lib.ptr(closest), _lib.ptr(uvw), _lib.ptr(ws),
                      wsb.value, ctypes.byref(n), self._s())
        return dist, face, closest, uvw, n.value

    def evaluate(self, P: torch.Tensor, band: float = 0.01, ref: torch.Tensor | None = None) -> dict:
        """Errors of the routed answers against the exact BVH on P (`ref`: the
        exact distances, when the caller has them already)."""
        d, face, _, _, n_exact = self.udf(P, band)
        if ref is None:
            ref = self.bvh.udf(P)[0]
        err = (d - ref).abs()
        far = face < 0
        out = {"n": int(P.shape[0]), "n_exact": n_exact, "frac_exact": n_exact / max(int(P.shape[0]), 1),
               "max_err": float(err.max()), "mean_err": float(err.mean()),
               "max_err_far": float(err[far].max()) if bool(far.any()) else 0.0,
               "mean_err_far": float(err[far].mean()) if bool(far.any()) else 0.0,
               "exact_match": bool(torch.equal(d[~far], ref[~far]))}
        # a query answered by the network although its true distance is inside the band
        out["missed_band"] = int(((ref < band) & far).sum())
        return out

    # ---------------------------------------------------------------- checkpoint
    def config(self) -> dict:
        return {"levels": L, "log2_table": LOG2T, "features": 2, "mlp": [32, 64, 64, 1], "lo": self.lo,
                "hi": self.hi, "n_min": 16, "n_max": 2048}

    def state_dict(self) -> dict:
        cfg = self.config()
        h = hashlib.sha256(json.dumps(cfg, sort_keys=True).encode()).hexdigest()
        return {"config": cfg, "config_hash": h, "params": self.params.cpu(), "step": self.step}

    def load_state_dict(self, sd: dict) -> None:
        cfg = self.config()
        h = hashlib.sha256(json.dumps(cfg, sort_keys=True).encode()).hexdigest()
        if sd.get("config_hash") != h:
            raise ValueError("surrogate checkpoint was made with a different configuration")
        self.params.copy_(sd["params"].to(self.device))
        self.step = int(sd.get("step", 0))


__all__ = ["UDFSurrogate"]

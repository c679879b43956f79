"""CPU oracle (test infrastructure only).

ctypes binding of ``liba3d_oracle.so`` — the C++ restatement of the
reference's algorithms in this folder.  Imported only by tests/, by
__graft_entry__.smoke() and by bench.py's cpu_baseline leg, as the checker;
the product (atom3d_amd) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_DIR = os.path.dirname(os.path.abspath(__file__))
# A3D_ORACLE_LIB: another build of the same sources (the sanitized one)
LIB_PATH = os.environ.get("A3D_ORACLE_LIB") or os.path.join(_DIR, "liba3d_oracle.so")
_lib = None

_f = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_d = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_l = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_u8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_i64 = ctypes.c_int64
_int = ctypes.c_int
_vp = ctypes.c_void_p

_SIGS = {
    "orc_num_threads": (_int, []),
    "orc_tri_query": (None, [_f, _f, _f, _f, _f]),
    "orc_tri_aabb_sat": (_int, [_f, _f, _f]),
    "orc_sah_build": (_i64, [_f, _i64, _i, _i64, _int, _f, _f, _vp]),
    "orc_bvh_udf": (None, [_f, _f, _f, _i64, _f, _i, _f, _f, _vp, _vp, _int]),
    "orc_brute_udf": (None, [_f, _i, _i64, _f, _i64, _f, _i, _f, _f, _int]),
    "orc_fp64_udf": (None, [_f, _i, _i64, _f, _i64, _d, _d, _int]),
    "orc_tri_closest_fp64": (None, [_d, _d, _d]),
    "orc_bvh_aabb": (_i64, [_f, _f, _f, _f, _i64, _u8, _i, _i, _i64, _int]),
    "orc_brute_aabb_hit": (None, [_f, _i, _i64, _f, _f, _i64, _u8, _vp, _int]),
    "orc_sat_clip": (None, [_f, _f, _f, _l, _l, _i64, _int, ctypes.c_float, _u8, _i, _f, _f, _f]),
    "orc_octree_traverse": (_i64, [_f, _f, _f, _int, _int, _l, _i64, _vp, _vp, _int, _int, _vp, _int]),
    "orc_leaf_brute_pairs": (None, [_f, _i64, _f, _f, _f, _i64, _i, _vp, _vp, _vp, _int]),
    "orc_ray_tri": (ctypes.c_float, [_f, _f, _f, _f]),
    "orc_brute_ray": (None, [_f, _i, _i64, _f, _f, _i64, ctypes.c_float, _u8, _f, _i, _f, _f, _int]),
    "orc_sah_ray": (None, [_f, _f, _f, _f, _i64, ctypes.c_float, _u8, _f, _i, _int, _vp]),
    "orc_brute_ray_all": (None, [_f, _i, _i64, _f, _f, _i64, ctypes.c_float, _l, _vp, _vp, _vp, _vp, _vp, _vp,
                                 _vp, _int]),
    "orc_crossings": (None, [_f, _i, _l, _i64, ctypes.c_float, _i]),
}


def build() -> str:
    subprocess.run(["make", "-C", _DIR, "-j8"], check=True, capture_output=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def _i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def num_threads() -> int:
    return lib().orc_num_threads()


def tri_query(tri9, p):
    """(d2, closest[3], uvw[3]) of the reference Triangle maths for one pair."""
    d2 = np.zeros(1, np.float32)
    c = np.zeros(3, np.float32)
    w = np.zeros(3, np.float32)
    lib().orc_tri_query(_f32(tri9).reshape(9), _f32(p).reshape(3), d2, c, w)
    return float(d2[0]), c, w


def tri_aabb_sat(tri9, bmin, bmax) -> bool:
    return bool(lib().orc_tri_aabb_sat(_f32(tri9).reshape(9), _f32(bmin).reshape(3), _f32(bmax).reshape(3)))


class SahBvh:
    """The reference's host SAH tree (nodes [Nn, 9] f32, tris [F, 10] f32).
    ``fallbacks`` counts the median splits: (after an empty SAH partition,
    with no axis to bin)."""

    def __init__(self, V, F, leaf: int = 4):
        V, F = _f32(V), _i32(F)
        nF = F.shape[0]
        nodes = np.zeros((2 * max(nF, 1) + 1, 9), np.float32)
        tris = np.zeros((max(nF, 1), 10), np.float32)
        st = np.zeros(2, np.int64)
        nn = lib().orc_sah_build(V.reshape(-1), V.shape[0], F.reshape(-1), nF, leaf, nodes.reshape(-1),
                                 tris.reshape(-1), st.ctypes.data)
        self.nodes = np.ascontiguousarray(nodes[:nn])
        self.tris = tris
        self.num_faces = nF
        self.fallbacks = (int(st[0]), int(st[1]))

    def order(self) -> np.ndarray:
        """Original face ids in the tree's triangle order (int32 [F])."""
        return np.ascontiguousarray(self.tris[:self.num_faces, 9]).view(np.int32).copy()

    def ranges(self):
        """{(start, end): box [6] float32} for every node: the triangle range
        a node covers in the tree's order and its box (min, max)."""
        nodes = self.nodes
        refs = nodes[:, 6:8].view(np.int32)
        out = {}
        if self.num_faces == 0:
            return out
        stack = [0]
        spans = {}
        order = []
        while stack:
            i = stack.pop()
            order.append(i)
            l, r = int(refs[i, 0]), int(refs[i, 1])
            if l < 0:
                spans[i] = (-l - 1, -r - 1)
            else:
                stack += [l, r]
        for i in reversed(order):  # children before parents
            l, r = int(refs[i, 0]), int(refs[i, 1])
            if l >= 0:
                spans[i] = (spans[l][0], spans[r][1])
            out[spans[i]] = nodes[i, :6].copy()
        return out

    def structure(self):
        """(leaves int64 [L, 2] — (start, end) triangle ranges, leaf boxes
        float32 [L, 6], internal-node boxes float32 [N, 6]) of this tree."""
        refs = self.nodes[:, 6:8].view(np.int32)
        leaf = refs[:, 0] < 0
        lv = np.stack([-refs[leaf, 0].astype(np.int64) - 1, -refs[leaf, 1].astype(np.int64) - 1], 1)
        return lv, self.nodes[leaf, :6].copy(), self.nodes[~leaf, :6].copy()

    def leaf_brute_pairs(self, qmin, qmax, threads: int = 0):
        """Walk-free reference box query on this tree: (hit [K] bool, box ids,
        face ids), pairs sorted by (box, face)."""
        qmin, qmax = _f32(qmin).reshape(-1, 3), _f32(qmax).reshape(-1, 3)
        K = qmin.shape[0]
        cnt = np.zeros(K, np.int32)
        args = (self.nodes.reshape(-1), self.nodes.shape[0], self.tris.reshape(-1), qmin.reshape(-1),
                qmax.reshape(-1), K, cnt)
        lib().orc_leaf_brute_pairs(*args, None, None, None, threads)
        off = np.zeros(K + 1, np.int64)
        np.cumsum(cnt, out=off[1:])
        pa = np.zeros(int(off[-1]), np.int32)
        pf = np.zeros(int(off[-1]), np.int32)
        if len(pa):
            lib().orc_leaf_brute_pairs(*args, off.ctypes.data, pa.ctypes.data, pf.ctypes.data, threads)
        return cnt > 0, pa, pf

    def udf(self, P, threads: int = 0, counts: bool = False):
        P = _f32(P).reshape(-1, 3)
        K = P.shape[0]
        d = np.zeros(K, np.float32)
        f = np.zeros(K, np.int32)
        c = np.zeros((K, 3), np.float32)
        w = np.zeros((K, 3), np.float32)
        nv = np.zeros(K, np.int32) if counts else None
        nt = np.zeros(K, np.int32) if counts else None
        lib().orc_bvh_udf(self.nodes.reshape(-1), self.tris.reshape(-1), P.reshape(-1), K, d, f, c.reshape(-1),
                          w.reshape(-1), nv.ctypes.data if counts else None, nt.ctypes.data if counts else None,
                          threads)
        return (d, f, c, w, nv, nt) if counts else (d, f, c, w)

    def ray(self, O, D, max_t: float = 1e10, threads: int = 0, work: bool = False):
        """The reference's nearest-hit walk on this tree: (hit, t, face[,
        (node fetches, triangle tests) summed over the rays])."""
        O, D = _f32(O).reshape(-1, 3), _f32(D).reshape(-1, 3)
        K = O.shape[0]
        hit = np.zeros(K, np.uint8)
        t = np.zeros(K, np.float32)
        f = np.zeros(K, np.int32)
        w = np.zeros(2, np.int64)
        lib().orc_sah_ray(self.nodes.reshape(-1), self.tris.reshape(-1), O.reshape(-1), D.reshape(-1), K,
                          float(max_t), hit, t, f, threads, w.ctypes.data)
        return (hit.astype(bool), t, f, (int(w[0]), int(w[1]))) if work else (hit.astype(bool), t, f)

    def aabb(self, qmin, qmax, canonical: bool = False, cap: int | None = None):
        """The reference's box walk on this tree: (hit [K] bool, box ids,
        face ids) with pairs in one thread's append order.  canonical adds
        the triangle's own AABB test (a tree-independent variant)."""
        qmin, qmax = _f32(qmin).reshape(-1, 3), _f32(qmax).reshape(-1, 3)
        K = qmin.shape[0]
        hit = np.zeros(K, np.uint8)
        cap = cap if cap is not None else max(64 * K, 1024)
        pa = np.zeros(cap, np.int32)
        pf = np.zeros(cap, np.int32)
        n = lib().orc_bvh_aabb(self.nodes.reshape(-1), self.tris.reshape(-1), qmin.reshape(-1), qmax.reshape(-1), K,
                               hit, pa, pf, cap, int(canonical))
        if n > cap:
            return self.aabb(qmin, qmax, canonical, cap=n)
        return hit.astype(bool), pa[:n], pf[:n]

    def octree(self, bounds, min_level: int, max_level: int, threads: int = 0, full: bool = False,
               work: bool = False, canonical: bool = False):
        """Octree level loop on this tree: (ijk int64 [n,3], tested per level,
        hits per level[, (node visits, triangle tests)]).  full walks every
        cube's whole overlap set as the reference does (no first-hit exit);
        canonical adds the triangle's own AABB test."""
        b = _f32(bounds).reshape(6)
        cap = 1 << 20
        while True:
            tested = np.zeros(max_level + 2, np.int64)
            hits = np.zeros(max_level + 2, np.int64)
            w = np.zeros(2, np.int64)
            out = np.zeros((cap, 3), np.int64)
            n = lib().orc_octree_traverse(self.nodes.reshape(-1), self.tris.reshape(-1), b, min_level, max_level,
                                          out.reshape(-1), cap, tested.ctypes.data, hits.ctypes.data, threads,
                                          int(full), w.ctypes.data, int(canonical))
            if n <= cap:
                return (out[:n], tested, hits, (int(w[0]), int(w[1]))) if work else (out[:n], tested, hits)
            cap = int(n)


def brute_udf(V, F, P, threads: int = 0):
    """Tree-independent closest face (ties to the lowest face index)."""
    V, F, P = _f32(V), _i32(F), _f32(P).reshape(-1, 3)
    K = P.shape[0]
    d = np.zeros(K, np.float32)
    f = np.zeros(K, np.int32)
    c = np.zeros((K, 3), np.float32)
    w = np.zeros((K, 3), np.float32)
    lib().orc_brute_udf(V.reshape(-1), F.reshape(-1), F.shape[0], P.reshape(-1), K, d, f, c.reshape(-1),
                        w.reshape(-1), threads)
    return d, f, c, w


def fp64_udf(V, F, P, threads: int = 0):
    """Exact FP64 distance and closest point (Ericson's Voronoi-region method)."""
    V, F, P = _f32(V), _i32(F), _f32(P).reshape(-1, 3)
    K = P.shape[0]
    d = np.zeros(K, np.float64)
    c = np.zeros((K, 3), np.float64)
    lib().orc_fp64_udf(V.reshape(-1), F.reshape(-1), F.shape[0], P.reshape(-1), K, d, c.reshape(-1), threads)
    return d, c


def tri_closest_fp64(tri9, p):
    out = np.zeros(3, np.float64)
    lib().orc_tri_closest_fp64(np.ascontiguousarray(tri9, np.float64).reshape(9),
                               np.ascontiguousarray(p, np.float64).reshape(3), out)
    return out


def brute_aabb_hit(V, F, qmin, qmax, threads: int = 0):
    V, F = _f32(V), _i32(F)
    qmin, qmax = _f32(qmin).reshape(-1, 3), _f32(qmax).reshape(-1, 3)
    K = qmin.shape[0]
    hit = np.zeros(K, np.uint8)
    cnt = np.zeros(K, np.int32)
    lib().orc_brute_aabb_hit(V.reshape(-1), F.reshape(-1), F.shape[0], qmin.reshape(-1), qmax.reshape(-1), K, hit,
                             cnt.ctypes.data, threads)
    return hit.astype(bool), cnt


def sat_clip(bmin, bmax, tris9, cand_a, cand_t, mode: int = 2, eps: float = 1e-8):
    """Reference polygon clip per (box, triangle) candidate:
    (hit [N] bool, counts [N], poly [N,8,3], centroids [N,3], areas [N])."""
    bmin, bmax, tris9 = _f32(bmin).reshape(-1, 3), _f32(bmax).reshape(-1, 3), _f32(tris9).reshape(-1, 9)
    ca = np.ascontiguousarray(cand_a, np.int64)
    ct = np.ascontiguousarray(cand_t, np.int64)
    N = len(ca)
    hit = np.zeros(N, np.uint8)
    cnt = np.zeros(N, np.int32)
    poly = np.zeros((N, 8, 3), np.float32)
    cen = np.zeros((N, 3), np.float32)
    area = np.zeros(N, np.float32)
    lib().orc_sat_clip(bmin.reshape(-1), bmax.reshape(-1), tris9.reshape(-1), ca, ct, N, mode, eps, hit, cnt,
                       poly.reshape(-1), cen.reshape(-1), area)
    return hit.astype(bool), cnt, poly, cen, area


def ray_tri(tri9, ro, rd):
    """(t, unit normal) of the reference's ray-triangle test; t = 1e10 when
    there is no hit."""
    n = np.zeros(3, np.float32)
    t = lib().orc_ray_tri(_f32(tri9).reshape(9), _f32(ro).reshape(3), _f32(rd).reshape(3), n)
    return float(np.float32(t)), n


def brute_ray(V, F, O, D, max_t: float = 1e10, threads: int = 0):
    """Tree-independent nearest hit (ties to the lowest face index):
    (hit, t, face, points, uvw)."""
    V, F, O, D = _f32(V), _i32(F), _f32(O).reshape(-1, 3), _f32(D).reshape(-1, 3)
    K = O.shape[0]
    hit = np.zeros(K, np.uint8)
    t = np.zeros(K, np.float32)
    f = np.zeros(K, np.int32)
    p = np.zeros((K, 3), np.float32)
    w = np.zeros((K, 3), np.float32)
    lib().orc_brute_ray(V.reshape(-1), F.reshape(-1), F.shape[0], O.reshape(-1), D.reshape(-1), K, float(max_t),
                        hit, t, f, p.reshape(-1), w.reshape(-1), threads)
    return hit.astype(bool), t, f, p, w


def brute_ray_all(V, F, O, D, max_t: float = 1e10, threads: int = 0):
    """Every hit of every ray, per ray in face order: dict of ray_idx, face,
    t, sign, points, uvw and offsets [K+1]."""
    V, F, O, D = _f32(V), _i32(F), _f32(O).reshape(-1, 3), _f32(D).reshape(-1, 3)
    K = O.shape[0]
    cnt = np.zeros(K, np.int64)
    args = (V.reshape(-1), F.reshape(-1), F.shape[0], O.reshape(-1), D.reshape(-1), K, float(max_t))
    lib().orc_brute_ray_all(*args, cnt, None, None, None, None, None, None, None, threads)
    off = np.zeros(K + 1, np.int64)
    np.cumsum(cnt, out=off[1:])
    H = int(off[-1])
    out = dict(ray_idx=np.zeros(H, np.int32), face=np.zeros(H, np.int32), t=np.zeros(H, np.float32),
               sign=np.zeros(H, np.int32), points=np.zeros((H, 3), np.float32), uvw=np.zeros((H, 3), np.float32))
    if H:
        lib().orc_brute_ray_all(*args, cnt, off.ctypes.data, out["ray_idx"].ctypes.data, out["face"].ctypes.data,
                                out["t"].ctypes.data, out["sign"].ctypes.data, out["points"].ctypes.data,
                                out["uvw"].ctypes.data, threads)
    out["offsets"] = off
    return out


def crossings(t, sign, offsets, coalesce_t: float = 1e-4):
    """Signed crossing count per ray from ray-major hit lists."""
    off = np.ascontiguousarray(offsets, np.int64)
    K = off.shape[0] - 1
    out = np.zeros(K, np.int32)
    lib().orc_crossings(_f32(t).reshape(-1), _i32(sign).reshape(-1), off, K, float(coalesce_t), out)
    return out

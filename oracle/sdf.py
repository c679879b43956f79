"""Oracle (test infrastructure only): the pseudo-normal SDF sign restated
in numpy float32.

Follows Atom3D atom3d/core/mesh_bvh.py:
  face normals                     :129-139  cross(v1-v0, v2-v0) / (|.| + 1e-8)
  vertex / edge pseudo-normals     :141-232  angle-weighted vertex normals,
                                             edge normal = sum of adjacent unit
                                             face normals, both / (|.| + 1e-12)
  sign                             :758-827  uvw < 1e-4: two zeros -> vertex
                                             normal (the non-zero corner), one
                                             zero -> edge normal (zero at corner
                                             c -> half-edge [1, 2, 0][c]);
                                             sign(dot(p - c, n)), 0 -> +1,
                                             face -1 -> +1
"""
from __future__ import annotations

import numpy as np


def _unit(x, eps):
    return (x / (np.linalg.norm(x, axis=-1, keepdims=True) + np.float32(eps))).astype(np.float32)


def pseudo_normals(V: np.ndarray, F: np.ndarray):
    """(face_n, vertex_pn, edge_pn, he_edge_idx) as the reference computes them."""
    V = np.asarray(V, np.float32)
    F = np.asarray(F, np.int64)
    v0, v1, v2 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    face_n = _unit(np.cross(v1 - v0, v2 - v0), 1e-8)

    def ang(a, b):
        c = (a * b).sum(-1) / (np.linalg.norm(a, axis=-1) * np.linalg.norm(b, axis=-1) + np.float32(1e-12))
        return np.arccos(np.clip(c, -1.0, 1.0)).astype(np.float32)

    angles = np.stack([ang(v1 - v0, v2 - v0), ang(v0 - v1, v2 - v1), ang(v0 - v2, v1 - v2)], -1)
    fn = _unit(np.cross(v1 - v0, v2 - v0), 1e-12)
    vn = np.zeros_like(V)
    for k in range(3):
        np.add.at(vn, F[:, k], angles[:, k:k + 1] * fn)
    vertex_pn = _unit(vn, 1e-12)
    he = np.stack([F, np.roll(F, -1, axis=1)], -1).reshape(-1, 2)
    key = np.minimum(he[:, 0], he[:, 1]) * len(V) + np.maximum(he[:, 0], he[:, 1])
    uniq, inv = np.unique(key, return_inverse=True)
    epn = np.zeros((len(uniq), 3), np.float32)
    np.add.at(epn, inv, fn[np.repeat(np.arange(len(F)), 3)])
    return face_n, vertex_pn, _unit(epn, 1e-12), inv.astype(np.int64)


def compute_sign(P, closest, face, uvw, face_n, vertex_pn, edge_pn, he_edge_idx, F):
    """(+-1 signs float32 [K], dot products float32 [K])."""
    P, closest, uvw = (np.asarray(x, np.float32) for x in (P, closest, uvw))
    face = np.asarray(face, np.int64)
    F = np.asarray(F, np.int64)
    K = len(P)
    signs = np.ones(K, np.float32)
    dots = np.zeros(K, np.float32)
    ok = face >= 0
    fi = face[ok]
    n = face_n[fi].copy()
    nz = uvw[ok] < np.float32(1e-4)
    cnt = nz.sum(1)
    vm = cnt == 2
    dom = np.argmax(~nz[vm], axis=1)
    n[vm] = vertex_pn[F[fi[vm], dom]]
    em = cnt == 1
    zi = np.argmax(nz[em], axis=1)
    n[em] = edge_pn[he_edge_idx[fi[em] * 3 + np.array([1, 2, 0])[zi]]]
    d = P[ok] - closest[ok]
    dot = (d[:, 0] * n[:, 0] + d[:, 1] * n[:, 1]) + d[:, 2] * n[:, 2]
    s = np.sign(dot).astype(np.float32)
    s[s == 0] = 1.0
    signs[ok] = s
    dots[ok] = dot
    return signs, dots

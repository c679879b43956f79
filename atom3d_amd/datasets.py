"""Deterministic synthetic workloads (meshes, query points, SDF grids).

All generators draw from ``numpy.random.default_rng`` with fixed seeds in
float64 and round once to float32, so the same arrays come out on every
machine.  They build the benchmark configurations:

* ``noisy_icosphere``: icosphere subdivided 5 times (10,242 vertices, 20,480
  faces), radius 0.5, each vertex moved radially by N(0, 0.01), seed 0;
* ``sphere_union``: 200 such icospheres (4,096,000 faces), centres uniform in
  [-0.8, 0.8]^3, radii log-uniform in [0.02, 0.2], uniformly random rotations,
  seed 1;
* ``uniform_points``: query points uniform in a box.
"""
from __future__ import annotations

import numpy as np
import torch

_T = (1.0 + 5.0 ** 0.5) / 2.0
_ICO_V = np.array([
    [-1, _T, 0], [1, _T, 0], [-1, -_T, 0], [1, -_T, 0],
    [0, -1, _T], [0, 1, _T], [0, -1, -_T], [0, 1, -_T],
    [_T, 0, -1], [_T, 0, 1], [-_T, 0, -1], [-_T, 0, 1],
], dtype=np.float64)
_ICO_F = np.array([
    [0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
    [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
    [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
    [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1],
], dtype=np.int64)


def icosphere(subdivisions: int = 5, radius: float = 1.0):
    """(vertices float64 [V,3], faces int64 [F,3]) of a geodesic sphere: the
    golden-ratio icosahedron, each triangle split in 4 through edge midpoints
    ``subdivisions`` times, every vertex projected to the sphere.  Faces are
    counter-clockwise seen from outside."""
    v = _ICO_V / np.linalg.norm(_ICO_V, axis=1, keepdims=True)
    f = _ICO_F.copy()
    for _ in range(subdivisions):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]], axis=0)
        lo, hi = np.minimum(e[:, 0], e[:, 1]), np.maximum(e[:, 0], e[:, 1])
        key = lo * len(v) + hi
        uniq, inv = np.unique(key, return_inverse=True)
        mids = (v[uniq // len(v)] + v[uniq % len(v)]) * 0.5
        mids /= np.linalg.norm(mids, axis=1, keepdims=True)
        m = inv.reshape(3, -1).T + len(v)  # midpoint of edges 01, 12, 20 per face
        v = np.concatenate([v, mids], axis=0)
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        m01, m12, m20 = m[:, 0], m[:, 1], m[:, 2]
        f = np.concatenate([
            np.stack([a, m01, m20], 1), np.stack([b, m12, m01], 1),
            np.stack([c, m20, m12], 1), np.stack([m01, m12, m20], 1),
        ], axis=0)
    return v * radius, f


def noisy_icosphere(subdivisions: int = 5, radius: float = 0.5, sigma: float = 0.01, seed: int = 0):
    """Icosphere with each vertex displaced radially by N(0, sigma)."""
    v, f = icosphere(subdivisions, 1.0)
    rng = np.random.default_rng(seed)
    r = radius + rng.normal(0.0, sigma, size=len(v))
    return torch.from_numpy((v * r[:, None]).astype(np.float32)), torch.from_numpy(f.astype(np.int32))


def _random_rotations(rng, n):
    # Uniform rotations from normalised Gaussian quaternions.
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], 1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], 1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], 1),
    ], 1)


def sphere_union(n_spheres: int = 200, subdivisions: int = 5, seed: int = 1, r_min: float = 0.02,
                 r_max: float = 0.2, extent: float = 0.8):
    """Concatenation of ``n_spheres`` rotated, scaled, translated icospheres."""
    v0, f0 = icosphere(subdivisions, 1.0)
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-extent, extent, size=(n_spheres, 3))
    radii = np.exp(rng.uniform(np.log(r_min), np.log(r_max), size=n_spheres))
    rots = _random_rotations(rng, n_spheres)
    vs = np.einsum("sij,vj->svi", rots, v0) * radii[:, None, None] + centres[:, None, :]
    fs = f0[None, :, :] + (np.arange(n_spheres) * len(v0))[:, None, None]
    return (torch.from_numpy(vs.reshape(-1, 3).astype(np.float32)),
            torch.from_numpy(fs.reshape(-1, 3).astype(np.int32)))


def uniform_points(n: int, lo: float = -1.0, hi: float = 1.0, seed: int = 0) -> torch.Tensor:
    rng = np.random.default_rng([seed, 1])
    return torch.from_numpy(rng.uniform(lo, hi, size=(n, 3)).astype(np.float32))


# ------------------------------------------------------------------ SDF shapes
def sdf_shapes(n_spheres: int = 32, n_tori: int = 16, seed: int = 2) -> dict:
    """Random spheres (r 0.05-0.3) and tori (R 0.1-0.3, r 0.02-0.08, axis x,
    y or z) with centres in [-0.7, 0.7]^3: the sparse-DMC workload."""
    rng = np.random.default_rng(seed)
    return {
        "sphere_c": rng.uniform(-0.7, 0.7, (n_spheres, 3)), "sphere_r": rng.uniform(0.05, 0.3, n_spheres),
        "torus_c": rng.uniform(-0.7, 0.7, (n_tori, 3)), "torus_R": rng.uniform(0.1, 0.3, n_tori),
        "torus_r": rng.uniform(0.02, 0.08, n_tori), "torus_axis": rng.integers(0, 3, n_tori),
    }


# two spheres and a torus: the small DMC test shape
SMALL_SHAPES = {
    "sphere_c": np.array([[-0.35, 0.0, 0.0], [0.4, 0.1, 0.05]]), "sphere_r": np.array([0.4, 0.3]),
    "torus_c": np.array([[0.0, -0.3, 0.35]]), "torus_R": np.array([0.3]), "torus_r": np.array([0.1]),
    "torus_axis": np.array([2]),
}


def shapes_sdf(P, shapes: dict | None = None):
    """Signed distance of the union (min) of the shapes at points P [K,3]
    (numpy or torch; torch stays on P's device and dtype)."""
    sh = SMALL_SHAPES if shapes is None else shapes
    tn = isinstance(P, torch.Tensor)
    if tn:
        T = {k: torch.as_tensor(v, device=P.device, dtype=P.dtype if k != "torus_axis" else torch.long)
             for k, v in sh.items()}
        norm = lambda x: torch.linalg.vector_norm(x, dim=-1)  # noqa: E731
        out = torch.full(P.shape[:1], float("inf"), device=P.device, dtype=P.dtype)
        mn = torch.minimum
    else:
        T = sh
        norm = lambda x: np.linalg.norm(x, axis=-1)  # noqa: E731
        out = np.full(P.shape[:1], np.inf)
        mn = np.minimum
    for c, r in zip(T["sphere_c"], T["sphere_r"]):
        out = mn(out, norm(P - c) - r)
    for c, R, r, ax in zip(T["torus_c"], T["torus_R"], T["torus_r"], T["torus_axis"]):
        d = P - c
        a = int(ax)
        u, v = [k for k in range(3) if k != a]
        q = ((d[:, u] ** 2 + d[:, v] ** 2) ** 0.5) - R
        out = mn(out, (q ** 2 + d[:, a] ** 2) ** 0.5 - r)
    return out


def dmc_band(res: int = 1024, shapes: dict | None = None, band: float = 0.8660254037844386, device="cuda",
             slab: int = 16):
    """Sparse narrow band for DMC on the res^3 grid over [-1,1]^3 (world =
    grid * 2 / res - 1): the cubes whose centre has |sdf| < band * h
    (default: half the cube diagonal, so every cube the surface crosses is
    kept), in lexicographic order.  Returns voxel_coords int32 [N,3], sdf
    float32 [M] at the unique corners (ascending linear corner id, evaluated
    in float64) and cube_idx int32 [N,8] in CUBE_CORNERS order."""
    sh = sdf_shapes() if shapes is None else shapes
    h = 2.0 / res
    g = torch.arange(res, device=device, dtype=torch.float64)
    cubes = []
    for x0 in range(0, res, slab):
        xs = torch.arange(x0, min(res, x0 + slab), device=device, dtype=torch.float64)
        X, Y, Z = torch.meshgrid(xs, g, g, indexing="ij")
        P = (torch.stack([X, Y, Z], -1).reshape(-1, 3) + 0.5) * h - 1.0
        idx = torch.nonzero(shapes_sdf(P, sh).abs() < band * h).squeeze(1)
        cubes.append(torch.stack([idx // (res * res) + x0, (idx // res) % res, idx % res], 1))
        del X, Y, Z, P
    cubes = torch.cat(cubes)
    R1 = res + 1
    off = torch.tensor([[(i >> 0) & 1, (i >> 1) & 1, (i >> 2) & 1] for i in range(8)], device=device)
    c = cubes[:, None, :] + off[None]
    lin = (c[..., 0] * R1 + c[..., 1]) * R1 + c[..., 2]
    used, inv = torch.unique(lin.reshape(-1), return_inverse=True)
    del c, lin
    sdf = torch.empty(used.shape[0], device=device, dtype=torch.float32)
    step = 1 << 24
    for s0 in range(0, used.shape[0], step):
        u = used[s0:s0 + step]
        Pc = torch.stack([u // (R1 * R1), (u // R1) % R1, u % R1], 1).double() * h - 1.0
        sdf[s0:s0 + step] = shapes_sdf(Pc, sh).float()
    return cubes.int(), sdf, inv.view(-1, 8).int()


# ------------------------------------------------------------------ uneven batch
def uneven_batch(n_meshes: int = 256, seed: int = 3, tri_min: int = 1000, tri_max: int = 200000,
                 q_min: int = 1 << 14, q_max: int = 1 << 20):
    """Meshes of log-uniform triangle counts (a noisy icosphere when one
    sphere of a subdivision level reaches the count, else a union of such
    spheres) paired with log-uniform query counts.  Returns a list of
    (V float32 [n,3], F int32 [m,3], n_queries)."""
    rng = np.random.default_rng(seed)
    out = []
    for m in range(n_meshes):
        target = float(np.exp(rng.uniform(np.log(tri_min), np.log(tri_max))))
        nq = int(np.exp(rng.uniform(np.log(q_min), np.log(q_max))))
        s = int(min(6, max(2, np.floor(np.log(target / 20.0) / np.log(4.0)))))
        k = max(1, int(round(target / (20 * 4 ** s))))
        if k == 1:
            V, F = noisy_icosphere(s, radius=0.5, sigma=0.01, seed=seed * 1000 + m)
        else:
            V, F = sphere_union(k, s, seed=seed * 1000 + m, r_min=0.05, r_max=0.3, extent=0.6)
        out.append((V, F, nq))
    return out


def batch_queries(meshes, seed: int = 3, pad: float = 0.1, device="cuda"):
    """Uniform queries in each mesh's bounding box padded by `pad` of its
    extent, concatenated: (P float32 [K,3] on device, counts list)."""
    g = torch.Generator(device=device).manual_seed(seed)
    parts, counts = [], []
    for V, _, nq in meshes:
        lo, hi = V.min(0)[0], V.max(0)[0]
        ext = hi - lo
        lo, hi = (lo - pad * ext).to(device), (hi + pad * ext).to(device)
        parts.append(torch.rand(nq, 3, generator=g, device=device) * (hi - lo) + lo)
        counts.append(nq)
    return torch.cat(parts).contiguous(), counts

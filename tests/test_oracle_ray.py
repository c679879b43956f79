"""Pins the CPU oracle's ray queries (the checker for the HIP ray kernels).

The ray-triangle test is checked against an FP64 plane intersection with an
FP64 inside test, against hand-made cases, and the all-faces nearest hit
against the reference's own walk over its SAH tree; crossing counts against
closed forms for closed spheres.
"""
import numpy as np
import pytest

import oracle
from atom3d_amd import datasets

TRI = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], np.float32)


def test_ray_tri_known_answers():
    t, n = oracle.ray_tri(TRI, [0.2, 0.2, 1.0], [0, 0, -1])
    assert t == 1.0 and np.array_equal(n, [0, 0, 1])
    t, _ = oracle.ray_tri(TRI, [0.2, 0.2, -3.0], [0, 0, 1])
    assert t == 3.0  # from below: a hit, two-sided
    assert oracle.ray_tri(TRI, [0.2, 0.2, 1.0], [0, 0, 1])[0] == 1e10  # behind the origin
    assert oracle.ray_tri(TRI, [0.2, 0.2, 1.0], [1, 0, 0])[0] == 1e10  # parallel to the plane
    assert oracle.ray_tri(TRI, [0.7, 0.7, 1.0], [0, 0, -1])[0] == 1e10  # outside (u + v = 1.4)
    # on an edge and on a vertex: inside within FLT_EPSILON
    assert oracle.ray_tri(TRI, [0.5, 0.5, 1.0], [0, 0, -1])[0] == 1.0
    assert oracle.ray_tri(TRI, [0.0, 0.0, 2.0], [0, 0, -1])[0] == 2.0
    assert oracle.ray_tri(TRI, [-1e-3, 0.5, 1.0], [0, 0, -1])[0] == 1e10
    # the origin on the plane: t = 0 is not a hit (t > 1e-8)
    assert oracle.ray_tri(TRI, [0.2, 0.2, 0.0], [0, 0, -1])[0] == 1e10


def _fp64_hit(tri, o, d):
    a, b, c = tri.reshape(3, 3).astype(np.float64)
    n = np.cross(b - a, c - a)
    den = n @ d
    if abs(den) < 1e-12:
        return None, None
    t = n @ (a - o) / den
    p = o + t * d
    # barycentrics of p
    v0, v1, v2 = b - a, c - a, p - a
    d00, d01, d11, d20, d21 = v0 @ v0, v0 @ v1, v1 @ v1, v2 @ v0, v2 @ v1
    den2 = d00 * d11 - d01 * d01
    v = (d11 * d20 - d01 * d21) / den2
    w = (d00 * d21 - d01 * d20) / den2
    return t, min(1 - v - w, v, w)


def test_ray_tri_against_fp64():
    rng = np.random.default_rng(0)
    checked = 0
    for _ in range(4000):
        tri = rng.uniform(-1, 1, 9).astype(np.float32)
        o = rng.uniform(-2, 2, 3).astype(np.float32)
        d = rng.normal(size=3)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        t_ref, margin = _fp64_hit(tri, o.astype(np.float64), d.astype(np.float64))
        t, _ = oracle.ray_tri(tri, o, d)
        if t_ref is None or abs(margin) < 1e-4 or abs(t_ref) < 1e-4:
            continue  # too close to an edge or to the origin to decide in FP32
        a, b, c = tri.reshape(3, 3).astype(np.float64)
        sin_g = abs(np.cross(b - a, c - a) @ d) / np.linalg.norm(np.cross(b - a, c - a))
        if sin_g < 1e-3:
            continue  # grazing: FP32 t is ill-conditioned
        should_hit = margin > 0 and t_ref > 0
        assert (t != 1e10) == should_hit, (tri, o, d, t, t_ref, margin)
        if should_hit:
            assert abs(t - t_ref) <= 1e-5 * max(1.0, abs(t_ref)) / sin_g, (t, t_ref)
        checked += 1
    assert checked > 2500


@pytest.fixture(scope="module")
def noisy():
    V, F = datasets.noisy_icosphere(subdivisions=3, seed=0)
    return V.numpy(), F.numpy()


def _rays(n, seed, inside=False):
    rng = np.random.default_rng(seed)
    O = rng.uniform(-0.2, 0.2, (n, 3)) if inside else rng.uniform(-1.0, 1.0, (n, 3))
    D = rng.normal(size=(n, 3))
    D /= np.linalg.norm(D, axis=1, keepdims=True)
    return O.astype(np.float32), D.astype(np.float32)


def test_brute_nearest_matches_reference_walk(noisy):
    V, F = noisy
    O, D = _rays(3000, 1)
    hit, t, f, p, w = oracle.brute_ray(V, F, O, D)
    sh, st, sf = oracle.SahBvh(V, F).ray(O, D)
    assert np.array_equal(hit, sh)
    assert np.array_equal(t, st)
    # the walk takes the first minimum in its order; the scan the lowest face
    same = f == sf
    assert same.mean() > 0.999, same.mean()
    for i in np.nonzero(~same)[0]:  # only exact ties may differ
        assert oracle.ray_tri(V[F[sf[i]]].reshape(9), O[i], D[i])[0] == t[i]
    assert hit.mean() > 0.1
    # outputs: point = o + d t, barycentrics sum to 1; misses keep max_t, -1
    np.testing.assert_allclose(p[hit], O[hit] + D[hit] * t[hit, None], rtol=0, atol=1e-6)
    np.testing.assert_allclose(w[hit].sum(1), 1.0, atol=1e-5)
    assert np.all(t[~hit] == np.float32(1e10)) and np.all(f[~hit] == -1)


def test_crossings_closed_sphere():
    V, F = datasets.icosphere(3, 0.5)
    V, F = V.astype(np.float32), F.astype(np.int32)
    for inside, expect in ((True, 1), (False, 0)):
        O, D = _rays(400, 2 + inside, inside=inside)
        if not inside:
            O = O * 3.0  # well outside the sphere
            O[np.linalg.norm(O, axis=1) < 0.7] *= 2.0
        h = oracle.brute_ray_all(V, F, O, D, max_t=1e6)
        c = oracle.crossings(h["t"], h["sign"], h["offsets"])
        assert np.all(c == expect), (inside, np.unique(c, return_counts=True))
        if inside:  # one exit per ray, leaving through a face turned along d
            assert np.all(np.diff(h["offsets"]) >= 1) and np.all(h["sign"] == 1)


def test_all_hits_consistent_with_nearest(noisy):
    V, F = noisy
    O, D = _rays(500, 5)
    hit, t, f, _, _ = oracle.brute_ray(V, F, O, D)
    h = oracle.brute_ray_all(V, F, O, D)
    off = h["offsets"]
    for r in range(len(O)):
        ts = h["t"][off[r]:off[r + 1]]
        fs = h["face"][off[r]:off[r + 1]]
        assert np.all(np.diff(fs) > 0)  # face order
        if ts.size:
            # the nearest of the filtered hits is never nearer than the nearest hit
            assert ts.min() >= t[r]
            if hit[r] and ts.min() == t[r]:
                assert fs[np.argmin(ts)] >= f[r]
    np.testing.assert_allclose(h["uvw"].sum(1), 1.0, atol=1e-5)
    assert set(np.unique(h["sign"])) <= {-1, 1}


def test_crossings_coalescing_rules():
    # one ray, hits at t = 1, 1 + 5e-5 (same group, signs +1 +1 -> +1),
    # 2 (alone, -1), 3, 3 (mixed group -1 +1 -> sum 0 -> 0)
    t = np.array([3.0, 1.0, 2.0, 1.00005, 3.0], np.float32)
    s = np.array([-1, 1, -1, 1, 1], np.int32)
    c = oracle.crossings(t, s, np.array([0, 5], np.int64), 1e-4)
    assert c[0] == 1 + (-1) + 0
    # mixed group of three with sum +1 -> parity 1 -> +1
    c = oracle.crossings(np.array([1, 1, 1], np.float32), np.array([1, -1, 1], np.int32), np.array([0, 3]), 1e-4)
    assert c[0] == 1

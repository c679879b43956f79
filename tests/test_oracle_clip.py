"""Pinning the oracle's polygon clip (no GPU): closed forms and FP64 clipping."""
import numpy as np
import pytest

import oracle


def _area_fp64(poly):
    if len(poly) < 3:
        return 0.0
    c = np.mean(poly, axis=0)
    return 0.5 * sum(np.linalg.norm(np.cross(poly[i] - c, poly[(i + 1) % len(poly)] - c)) for i in range(len(poly)))


def _clip_fp64(tri, lo, hi):
    poly = [np.asarray(v, np.float64) for v in np.asarray(tri, np.float64).reshape(3, 3)]
    for axis in range(3):
        for sgn, bound in ((-1, hi[axis]), (1, lo[axis])):
            out, n = [], len(poly)
            for i in range(n):
                a, b = poly[i], poly[(i + 1) % n]
                da, db = sgn * (a[axis] - bound), sgn * (b[axis] - bound)
                if da >= 0 and db >= 0:
                    out.append(b)
                elif (da >= 0) != (db >= 0):
                    out.append(a + da / (da - db) * (b - a))
                    if db >= 0:
                        out.append(b)
            poly = out
            if not poly:
                return []
    return poly


def test_triangle_inside_box_is_unchanged():
    tri = np.array([[0.1, 0.1, 0.1], [0.5, 0.2, 0.3], [0.2, 0.6, 0.4]], np.float32)
    hit, cnt, poly, cen, area = oracle.sat_clip([[0, 0, 0]], [[1, 1, 1]], tri.reshape(1, 9), [0], [0], mode=2)
    assert hit[0] and cnt[0] == 3
    np.testing.assert_array_equal(poly[0, :3], tri)          # six full passes rotate back
    assert np.all(poly[0, 3:] == 0)
    a = 0.5 * np.linalg.norm(np.cross(tri[1] - tri[0], tri[2] - tri[0]))
    assert area[0] == pytest.approx(a, rel=1e-5)
    np.testing.assert_allclose(cen[0], tri.mean(0), atol=1e-6)


def test_box_through_large_triangle_is_a_rectangle():
    tri = np.array([[-5, -5, 0.25], [5, -5, 0.25], [0, 5, 0.25]], np.float32)   # plane z = 0.25
    hit, cnt, poly, cen, area = oracle.sat_clip([[0, 0, 0]], [[1, 0.5, 1]], tri.reshape(1, 9), [0], [0], mode=2)
    assert hit[0] and cnt[0] == 4
    assert area[0] == pytest.approx(0.5, rel=1e-6)
    np.testing.assert_allclose(cen[0], [0.5, 0.25, 0.25], atol=1e-6)


def test_miss_and_modes():
    tri = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], np.float32)
    hit, cnt, poly, cen, area = oracle.sat_clip([[2, 2, 2]], [[3, 3, 3]], tri.reshape(1, 9), [0], [0], mode=2)
    assert not hit[0] and cnt[0] == 0 and area[0] == 0 and np.all(cen == 0) and np.all(poly == 0)


def test_random_against_fp64_clipping():
    rng = np.random.default_rng(0)
    n_cmp = 0
    for _ in range(3000):
        tri = rng.uniform(-1, 1, 9).astype(np.float32)
        c = rng.uniform(-0.5, 0.5, 3)
        h = rng.uniform(0.05, 0.6, 3)
        lo, hi = (c - h).astype(np.float32), (c + h).astype(np.float32)
        hit, cnt, poly, cen, area = oracle.sat_clip(lo[None], hi[None], tri[None], [0], [0], mode=2)
        ref = _clip_fp64(tri, lo.astype(np.float64), hi.astype(np.float64))
        ra = _area_fp64(np.array(ref)) if ref else 0.0
        if ra > 1e-3:  # clearly intersecting: same polygon up to FP32 rounding
            assert hit[0]
            assert area[0] == pytest.approx(ra, rel=1e-3, abs=1e-5)
            V = tri.reshape(3, 3).astype(np.float64)
            nrm = np.cross(V[1] - V[0], V[2] - V[0])
            nrm /= np.linalg.norm(nrm)
            assert abs(np.dot(cen[0] - V[0], nrm)) < 1e-4          # centroid on the triangle's plane
            n_cmp += 1
        if not ref:
            # FP64 says empty: the FP32 clip may still keep a sliver within eps
            assert area[0] < 1e-4
    assert n_cmp > 500

"""Pinning the CPU oracle (no GPU).

The oracle restates the reference's FP32 geometry; these tests check it
against things the restatement cannot share a mistake with: closed forms,
an independent FP64 closest-point method, the variational characterisation
of a projection onto a convex set, exact FP64 polygon clipping for the SAT,
and brute force over every face / every cube.
"""
import numpy as np
import pytest

import oracle
from atom3d_amd import datasets

TRI = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], np.float32)


@pytest.mark.parametrize("p,d2,cp,uvw", [
    ((0.2, 0.2, 0.5), 0.25, (0.2, 0.2, 0.0), (0.6, 0.2, 0.2)),    # face interior (plane branch)
    ((-1.0, -1.0, 0.0), 2.0, (0.0, 0.0, 0.0), (1.0, 0.0, 0.0)),   # vertex a region
    ((0.5, -1.0, 0.0), 1.0, (0.5, 0.0, 0.0), (0.5, 0.5, 0.0)),    # edge ab region
    ((1.0, 1.0, 1.0), 1.5, (0.5, 0.5, 0.0), (0.0, 0.5, 0.5)),     # edge bc region, off plane
    ((2.0, -1.0, 3.0), 11.0, (1.0, 0.0, 0.0), (0.0, 1.0, 0.0)),   # vertex b region
])
def test_triangle_closed_forms(p, d2, cp, uvw):
    g, c, w = oracle.tri_query(TRI, np.array(p, np.float32))
    assert g == pytest.approx(d2, rel=1e-6)
    np.testing.assert_allclose(c, cp, atol=1e-6)
    np.testing.assert_allclose(w, uvw, atol=1e-6)


def test_fp64_closest_point_is_the_projection():
    """q = proj_T(p) iff q in T and (p - q).(x - q) <= 0 for every x in T; as
    the inequality is linear in x it suffices to check the three vertices."""
    rng = np.random.default_rng(0)
    for _ in range(2000):
        t = rng.normal(size=9)
        if rng.random() < 0.1:  # thin triangles
            t[6:9] = t[0:3] + 1e-3 * rng.normal(size=3) + (t[3:6] - t[0:3]) * rng.random()
        p = rng.normal(size=3) * 2
        q = oracle.tri_closest_fp64(t, p)
        V = t.reshape(3, 3)
        scale = np.abs(V).max() + np.abs(p).max()
        for x in V:
            assert np.dot(p - q, x - q) <= 1e-9 * scale * scale
        # q is a convex combination of the vertices
        A = np.vstack([V.T, np.ones(3)])
        lam = np.linalg.lstsq(A, np.append(q, 1.0), rcond=None)[0]
        if np.linalg.matrix_rank(A) == 4 or np.linalg.cond(A) < 1e8:
            assert lam.min() >= -1e-7


def test_fp32_distance_vs_fp64():
    rng = np.random.default_rng(1)
    for _ in range(3000):
        t = rng.uniform(-1, 1, 9).astype(np.float32)
        p = rng.uniform(-2, 2, 3).astype(np.float32)
        d2, c, w = oracle.tri_query(t, p)
        q = oracle.tri_closest_fp64(t.astype(np.float64), p.astype(np.float64))
        d64 = np.linalg.norm(p.astype(np.float64) - q)
        assert abs(np.sqrt(d2) - d64) <= 2e-6 * (1 + d64)
        # closest point consistent with the returned distance
        assert abs(np.linalg.norm(c.astype(np.float64) - p) - np.sqrt(d2)) <= 2e-6 * (1 + d64)
        # barycentrics reproduce the closest point and sum to one
        V = t.reshape(3, 3).astype(np.float64)
        area = 0.5 * np.linalg.norm(np.cross(V[1] - V[0], V[2] - V[0]))
        quality = 4 * np.sqrt(3) * area / sum(np.sum((V[i] - V[(i + 1) % 3]) ** 2) for i in range(3))
        if quality > 0.2:  # well-shaped: barycentrics are well conditioned
            np.testing.assert_allclose(w @ V, c, atol=1e-4)
        assert abs(w.sum() - 1.0) < 1e-5


def _mesh(sub=3, seed=0):
    V, F = datasets.noisy_icosphere(subdivisions=sub, seed=seed)
    return V.numpy(), F.numpy()


def test_brute_udf_lowest_index_rule_and_fp64():
    V, F = _mesh(2, 3)
    P = datasets.uniform_points(300, seed=3).numpy()
    d, f, c, w = oracle.brute_udf(V, F, P)
    for q in range(len(P)):
        all_d2 = np.array([oracle.tri_query(V[F[k]].reshape(9), P[q])[0] for k in range(len(F))], np.float32)
        m = all_d2.min()
        assert f[q] == np.nonzero(all_d2 == m)[0][0]
        assert d[q] == np.sqrt(np.float32(m))
    dd, cc = oracle.fp64_udf(V, F, P)
    assert np.max(np.abs(d - dd)) < 2e-6


def test_sah_tree_structure_and_walk():
    V, F = _mesh(4, 1)
    sah = oracle.SahBvh(V, F)
    nodes, tris = sah.nodes, sah.tris
    ids = tris[:, 9].view(np.int32)
    assert np.array_equal(np.sort(ids), np.arange(len(F)))
    left, right = nodes[:, 6].view(np.int32), nodes[:, 7].view(np.int32)
    leaf = left < 0
    s, e = -left[leaf] - 1, -right[leaf] - 1
    assert np.all(e - s <= 4) and np.all(e > s)
    cover = np.zeros(len(F), int)
    for a, b in zip(s, e):
        cover[a:b] += 1
    assert np.all(cover == 1)
    # every leaf box contains its triangles
    for n in np.nonzero(leaf)[0]:
        a, b = -left[n] - 1, -right[n] - 1
        t = tris[a:b, :9].reshape(-1, 3)
        assert np.all(t.min(0) >= nodes[n, 0:3]) and np.all(t.max(0) <= nodes[n, 3:6])
    # the walk finds the FP32 minimum up to the reference's 1-2 ulp pruning slip
    P = datasets.uniform_points(3000, seed=4).numpy()
    d, f, c, w = sah.udf(P)
    bd, bf, bc, bw = oracle.brute_udf(V, F, P)
    assert np.all(bd <= d) and np.all(d - bd <= 2.5e-7 * np.maximum(d, 1e-3))


def _clip_fp64(tri, lo, hi):
    """Exact-arithmetic-ish (FP64) Sutherland-Hodgman clip of a triangle by a
    box; returns the clipped polygon's vertices."""
    poly = [np.asarray(v, np.float64) for v in tri.reshape(3, 3)]
    for axis in range(3):
        for sgn, bound in ((1, lo[axis]), (-1, hi[axis])):
            out = []
            n = len(poly)
            for i in range(n):
                a, b = poly[i], poly[(i + 1) % n]
                da, db = sgn * (a[axis] - bound), sgn * (b[axis] - bound)
                if da >= 0:
                    out.append(a)
                if (da >= 0) != (db >= 0):
                    t = da / (da - db)
                    out.append(a + t * (b - a))
            poly = out
            if not poly:
                return poly
    return poly


def test_sat_against_exact_clipping():
    rng = np.random.default_rng(2)
    agree = 0
    for _ in range(4000):
        t = rng.uniform(-1, 1, 9).astype(np.float32) * rng.uniform(0.05, 1)
        c = rng.uniform(-0.8, 0.8, 3)
        h = rng.uniform(0.01, 0.5, 3)
        lo, hi = (c - h).astype(np.float32), (c + h).astype(np.float32)
        sat = oracle.tri_aabb_sat(t, lo, hi)
        inter = len(_clip_fp64(t, lo.astype(np.float64), hi.astype(np.float64))) > 0
        # separation margin: distance of the clipped result from empty
        big = len(_clip_fp64(t, lo.astype(np.float64) + 1e-4, hi.astype(np.float64) - 1e-4)) > 0
        small = len(_clip_fp64(t, lo.astype(np.float64) - 1e-4, hi.astype(np.float64) + 1e-4)) > 0
        if big:
            assert sat, "box clearly intersects the triangle but SAT rejects it"
        if not small:
            assert not sat, "box is clearly apart from the triangle but SAT accepts it"
        agree += sat == inter
    assert agree >= 3990


def test_sat_touching_cases():
    tri = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], np.float32)
    # box resting on the triangle's plane: touching counts (inclusive, eps)
    assert oracle.tri_aabb_sat(tri, [0.1, 0.1, 0.0], [0.2, 0.2, 0.1])
    assert oracle.tri_aabb_sat(tri, [0.1, 0.1, 5e-7], [0.2, 0.2, 0.1])      # within eps 1e-6
    assert not oracle.tri_aabb_sat(tri, [0.1, 0.1, 1e-5], [0.2, 0.2, 0.1])  # beyond eps
    # box beyond the hypotenuse, touching its corner region only by distance
    assert not oracle.tri_aabb_sat(tri, [0.6, 0.6, -0.1], [0.7, 0.7, 0.1])
    assert oracle.tri_aabb_sat(tri, [0.45, 0.45, -0.1], [0.7, 0.7, 0.1])


def _octree_order_key(ijk, min_level, max_level):
    keys = []
    for v in ijk:
        top = [int(x) >> (max_level - min_level) for x in v]
        r = 1 << min_level
        k = [(top[0] * r + top[1]) * r + top[2]]
        for lv in range(min_level + 1, max_level + 1):
            sh = max_level - lv
            k.append(((v[0] >> sh) & 1) | (((v[1] >> sh) & 1) << 1) | (((v[2] >> sh) & 1) << 2))
        keys.append(tuple(k))
    return keys


@pytest.mark.parametrize("min_level,max_level", [(1, 4), (2, 5), (3, 3)])
def test_octree_oracle_vs_brute_force(min_level, max_level):
    """The octree's kept set equals the walk-free leaf-box statement of the
    reference's box query on every cube of max_level (a cube whose parent
    was dropped has no hit either: node boxes contain their children's, and
    the parent cube contains the child cube)."""
    V, F = _mesh(2, 5)
    sah = oracle.SahBvh(V, F)
    got, tested, hits = sah.octree([-1, -1, -1, 1, 1, 1], min_level, max_level)
    r = 1 << max_level
    g = np.stack(np.meshgrid(np.arange(r), np.arange(r), np.arange(r), indexing="ij"), -1).reshape(-1, 3)
    cell = np.float32(2.0) / np.float32(r)
    lo = (g.astype(np.float32) * cell + np.float32(-1)).astype(np.float32)
    hi = (lo + cell).astype(np.float32)
    hit, _, _ = sah.leaf_brute_pairs(lo, hi)
    brute = g[hit]
    assert {tuple(x) for x in got} == {tuple(x) for x in brute}
    keys = _octree_order_key(got, min_level, max_level)
    assert keys == sorted(keys), "octree output is not in parent-major x CUBE_CORNERS order"
    assert hits[max_level] == len(got)
    full, _, _ = sah.octree([-1, -1, -1, 1, 1, 1], min_level, max_level, full=True)
    np.testing.assert_array_equal(full, got)


def test_reference_walk_equals_leaf_brute_force():
    """The reference walk (node-box overlap descent, SAT at the leaves) keeps
    exactly the pairs whose leaf box overlaps and whose SAT passes, found by
    scanning every leaf; its canonical variant (the triangle's own box
    overlapping as well) is a subset, equal to the all-faces brute force."""
    V, F = _mesh(3, 6)
    sah = oracle.SahBvh(V, F)
    rng = np.random.default_rng(9)
    c = rng.uniform(-1, 1, (1000, 3))
    h = rng.uniform(0.005, 0.1, (1000, 3))
    lo, hi = (c - h).astype(np.float32), (c + h).astype(np.float32)
    h0, a0, f0 = sah.aabb(lo, hi)
    hb, ab, fb = sah.leaf_brute_pairs(lo, hi)
    np.testing.assert_array_equal(h0, hb)
    o = np.lexsort((f0, a0))
    np.testing.assert_array_equal(a0[o], ab)
    np.testing.assert_array_equal(f0[o], fb)
    h1, a1, f1 = sah.aabb(lo, hi, canonical=True)
    s1, s0 = set(zip(a1, f1)), set(zip(a0, f0))
    assert s1 <= s0
    hc, cc = oracle.brute_aabb_hit(V, F, lo, hi)
    np.testing.assert_array_equal(h1, hc)
    np.testing.assert_array_equal(np.bincount(a1, minlength=1000), cc)


def test_reference_walk_keeps_eps_band_faces():
    """A box 5e-7 beyond a face's own box along x (within the SAT slack):
    the face is kept iff its leaf box reaches the query box."""
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    far = np.array([[5, 5, 5], [6, 5, 5], [5, 6, 5]], np.float32)
    wide = np.array([[0, 2, 0], [2.0, 2, 0], [0, 2, 0.1]], np.float32)  # extends the leaf box past x = 1
    V = np.concatenate([tri, wide, far]).astype(np.float32)
    F = np.arange(9, dtype=np.int32).reshape(3, 3)
    lo = np.array([[1.0 + 5e-7, 0.0, -0.1]], np.float32)
    hi = np.array([[1.5, 0.5, 0.1]], np.float32)
    assert oracle.tri_aabb_sat(tri.reshape(-1), lo[0], hi[0])
    sah = oracle.SahBvh(V, F, leaf=4)  # one leaf holds all three faces
    hit, pa, pf = sah.aabb(lo, hi)
    assert hit[0] and 0 in set(pf.tolist())
    _, _, cf = sah.aabb(lo, hi, canonical=True)
    assert 0 not in set(cf.tolist())
    sah1 = oracle.SahBvh(V, F, leaf=1)  # the face alone in its leaf: its own box misses
    _, _, pf1 = sah1.aabb(lo, hi)
    assert 0 not in set(pf1.tolist())

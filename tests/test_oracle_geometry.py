"""Pins the oracle's intersection and BVH.

- Moller-Trumbore vs axis-aligned closed forms.
- A quad hit (the reference's quad test, restated below from
  backends/metal/plugin.mm:157-195) iff one of its two triangles is hit, at
  the same t.
- BVH closest hit == brute force over all triangles (same triangle, same t).
- BVH invariants: every triangle in exactly one leaf, leaves <= 4, boxes
  contain their children and triangles.
"""
import numpy as np
import pytest

from oracle import oracle as O
from quasi_amd import scenes


def quad_hit_reference(o, d, q0, u, v, tmin=1e-3, tmax=1e30):
    """plugin.mm:157-195 in float64: plane hit, alpha = w.(p x v), beta = w.(u x p)."""
    n = np.cross(u, v)
    area_sq = n @ n
    if area_sq < 1e-8:
        return None
    normal = n / np.sqrt(area_sq)
    dd = normal @ q0
    denom = normal @ d
    if abs(denom) < 1e-8:
        return None
    t = (dd - normal @ o) / denom
    if t < tmin or t > tmax:
        return None
    p = o + d * t - q0
    w = n / area_sq
    alpha = w @ np.cross(p, v)
    beta = w @ np.cross(u, p)
    if alpha < 0 or alpha > 1 or beta < 0 or beta > 1:
        return None
    return t, alpha, beta


def test_triangle_axis_aligned_closed_form():
    p0, p1, p2 = (-1, 0, -1), (1, 0, -1), (1, 0, 1)
    assert O.intersect_tri((0.5, 5, -0.2), (0, -1, 0), p0, p1, p2) == pytest.approx(5.0, abs=0)
    assert O.intersect_tri((0.5, -2, -0.2), (0, 1, 0), p0, p1, p2) == pytest.approx(2.0, abs=0)  # two-sided
    assert O.intersect_tri((-0.5, 5, 0.2), (0, -1, 0), p0, p1, p2) is None                         # other half
    assert O.intersect_tri((0.5, 5, -0.2), (1, 0, 0), p0, p1, p2) is None                           # parallel
    assert O.intersect_tri((0.5, 5, -0.2), (0, -1, 0), p0, p1, p2, tmax=4.999) is None              # beyond tmax
    assert O.intersect_tri((0.5, 5, -0.2), (0, -1, 0), p0, p1, p2, tmax=5.0) == 5.0                 # inclusive


def test_triangle_edges_are_inclusive():
    p0, p1, p2 = (0, 0, 0), (1, 0, 0), (0, 1, 0)
    # vertex and edge midpoints
    for x, y in [(0, 0), (1, 0), (0, 1), (0.5, 0), (0, 0.5), (0.5, 0.5)]:
        assert O.intersect_tri((x, y, 1), (0, 0, -1), p0, p1, p2) == 1.0


def test_quad_iff_two_triangles():
    rng = np.random.default_rng(5)
    checked = 0
    for _ in range(3000):
        q0 = rng.uniform(-1, 1, 3)
        u = rng.uniform(-1, 1, 3)
        v = rng.uniform(-1, 1, 3)
        o = rng.uniform(-3, 3, 3)
        target = q0 + rng.uniform(-0.2, 1.2) * u + rng.uniform(-0.2, 1.2) * v
        d = target - o
        d /= np.linalg.norm(d)
        qh = quad_hit_reference(o, d, q0, u, v)
        if qh is not None:
            t, a, b = qh
            # skip rays within 1e-7 of an edge or the diagonal where both
            # rules are ill-conditioned
            if min(a, 1 - a, b, 1 - b, abs(a - b)) < 1e-7:
                continue
        t1 = O.intersect_tri(o, d, q0, q0 + u, q0 + u + v)
        t2 = O.intersect_tri(o, d, q0, q0 + u + v, q0 + v)
        hits = [t for t in (t1, t2) if t is not None]
        if qh is None:
            assert not hits
        else:
            assert len(hits) == 1
            assert hits[0] == pytest.approx(qh[0], rel=1e-10)
            checked += 1
    assert checked > 300


@pytest.fixture(scope="module")
def soup_scene():
    m = scenes.soup(30000, seed=4)
    return m, O.Scene(m)


def test_bvh_closest_hit_equals_brute_force(soup_scene):
    m, s = soup_scene
    rng = np.random.default_rng(9)
    n = 1500
    o = rng.uniform(-1.5, 1.5, size=(n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    tb, t_b = s.closest_hits(o, d, use_bvh=True)
    tf, t_f = s.closest_hits(o, d, use_bvh=False)
    assert np.array_equal(tb, tf)
    assert np.array_equal(t_b, t_f)
    assert (tb != 0xFFFFFFFF).sum() > n // 10


def test_bvh_invariants(soup_scene):
    m, s = soup_scene
    lo, hi, first, count, prims = s.bvh()
    info = s.info()
    assert info["max_leaf"] <= 4
    leaves = count > 0
    # every triangle referenced by exactly one leaf
    covered = np.zeros(m.num_triangles, dtype=np.int64)
    for f, c in zip(first[leaves], count[leaves]):
        covered[prims[f:f + c]] += 1
    assert (covered == 1).all()
    tri = m.triangles()
    tlo, thi = tri.min(1), tri.max(1)
    for i in np.nonzero(leaves)[0][:2000]:
        p = prims[first[i]:first[i] + count[i]]
        assert (lo[i] <= tlo[p]).all() and (hi[i] >= thi[p]).all()
    internal = np.nonzero(~leaves)[0]
    for i in internal[:2000]:
        for c in (first[i], first[i] + 1):
            assert (lo[i] <= lo[c]).all() and (hi[i] >= hi[c]).all()


def test_bvh_small_scene_reference_cornell():
    s = O.Scene(scenes.cornell_ref())
    info = s.info()
    assert info["lights"] == 2 and info["max_leaf"] <= 4

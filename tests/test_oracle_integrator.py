"""Pins the oracle's sampling and light transport with closed forms.

- cosine_sample (plugin.mm:111-125): cos(theta) = sqrt(1 - r2) exactly, unit
  length, upper hemisphere, E[d] = (2/3) n, and the density cos(theta)/pi
  (Kolmogorov-Smirnov on cos(theta) and on the azimuth).
- Reference bounce loop (plugin.mm:223-276): furnace closed form.  In a closed
  box whose walls all have albedo a and emission E <= 0.1 (so no wall counts
  as an emitter), every path sees E * sum_{k<D} a^k exactly.
- Both integrators: direct lighting of a floor point under a parallel square
  emitter equals albedo * Le * F, with F the analytic point-to-rectangle view
  factor (a formula independent of any path tracing).
- NEE and the reference loop have the same expectation (batch-means test).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from quasi_amd import scenes
from quasi_amd.scenes import F32, Mesh, _Builder


def test_cosine_sample_cos_theta_is_sqrt_one_minus_r2():
    rng = np.random.default_rng(2)
    for _ in range(200):
        n = rng.normal(size=3)
        st = int(rng.integers(0, 2**32))
        d = O.cosine_sample(n, st)
        r1, r2 = O.random_floats(st, 2)
        nn = n / np.linalg.norm(n)
        assert np.linalg.norm(d) == pytest.approx(1.0, abs=1e-14)
        assert d @ nn == pytest.approx(math.sqrt(1 - float(r2)), abs=1e-12)


@pytest.mark.parametrize("n", [(0, 1, 0), (0.95, 0.1, 0.0), (-0.3, -0.2, 0.9)])
def test_cosine_sample_moments(n):
    n = np.array(n, dtype=np.float64)
    nn = n / np.linalg.norm(n)
    states = np.random.default_rng(7).integers(0, 2**32, size=20000)
    ds = np.array([O.cosine_sample(n, int(s)) for s in states])
    c = ds @ nn
    assert (c >= 0).all()
    assert c.mean() == pytest.approx(2 / 3, abs=0.006)        # E[cos] = 2/3
    assert (c ** 2).mean() == pytest.approx(0.5, abs=0.006)   # E[cos^2] = 1/2
    tang = ds - np.outer(c, nn)
    assert np.abs(tang.mean(0)).max() < 0.01                   # no azimuthal bias


@pytest.mark.parametrize("n", [(0, 1, 0), (0.95, 0.1, 0.0), (-0.3, -0.2, 0.9)])
def test_cosine_sample_pdf_ks(n):
    """Directions follow the density cos(theta)/pi: P(cos theta <= c) = c^2 and the
    azimuth, measured in a tangent basis built here independently of the
    sampler's, is uniform (Kolmogorov-Smirnov on both)."""
    from scipy import stats
    n = np.array(n, dtype=np.float64)
    nn = n / np.linalg.norm(n)
    states = np.random.default_rng(11).integers(0, 2**32, size=20000)
    ds = np.array([O.cosine_sample(n, int(s)) for s in states])
    c = np.clip(ds @ nn, 0.0, 1.0)
    assert stats.kstest(c, lambda x: np.clip(x, 0, 1) ** 2).pvalue > 1e-3
    t1 = np.cross(nn, [0.0, 0.0, 1.0] if abs(nn[2]) < 0.5 else [1.0, 0.0, 0.0])
    t1 /= np.linalg.norm(t1)
    t2 = np.cross(nn, t1)
    phi = np.arctan2(ds @ t2, ds @ t1) % (2 * np.pi)
    assert stats.kstest(phi / (2 * np.pi), "uniform").pvalue > 1e-3
    # a sampler with the wrong density (uniform over the hemisphere) fails the same test
    assert stats.kstest(np.sqrt(c), lambda x: np.clip(x, 0, 1) ** 2).pvalue < 1e-6


def closed_box(albedo, emission, half=1.0):
    b = _Builder()
    h = half
    faces = [((-h, -h, -h), (2 * h, 0, 0), (0, 0, 2 * h)), ((-h, h, h), (2 * h, 0, 0), (0, 0, -2 * h)),
             ((-h, -h, -h), (2 * h, 0, 0), (0, 2 * h, 0)), ((-h, -h, h), (0, 2 * h, 0), (2 * h, 0, 0)),
             ((-h, -h, h), (0, 0, -2 * h), (0, 2 * h, 0)), ((h, -h, -h), (0, 0, 2 * h), (0, 2 * h, 0))]
    for o, u, v in faces:
        b.quad(o, u, v, (albedo,) * 3, (emission,) * 3)
    cam = dict(position=(0.0, 0.0, 0.0), target=(0.0, 0.0, -1.0), up=(0.0, 1.0, 0.0), fov=60.0)
    return b.mesh(cam, "furnace")


@pytest.mark.parametrize("depth", [1, 3, 5])
def test_furnace_closed_form_reference_loop(depth):
    a, e = 0.5, 0.1
    s = O.Scene(closed_box(a, e))
    e = float(np.float32(e))   # the scene stores float32 radiance
    r = s.render(16, 16, 8, max_depth=depth, rr_start=depth, integrator=0, rng_mode=1, seed=3)
    expect = e * sum(a ** k for k in range(depth))
    np.testing.assert_allclose(r.mean()[..., :3], expect, rtol=1e-12)
    assert r.counts.camera_rays == 16 * 16 * 8
    assert r.counts.extension_rays == 16 * 16 * 8 * (depth - 1)


def test_furnace_with_russian_roulette_is_unbiased():
    a, e = 0.5, 0.1
    s = O.Scene(closed_box(a, e))
    r = s.render(32, 32, 64, max_depth=8, rr_start=3, integrator=0, rng_mode=1, seed=4)
    expect = e * sum(a ** k for k in range(8))
    assert r.mean()[..., :3].mean() == pytest.approx(expect, rel=0.01)


def view_factor_point_to_centered_square(w, h):
    """Differential area to a parallel square of half-side w centred at height h:
    four corner rectangles, F_corner = 1/(2 pi) [A/sqrt(1+A^2) atan(B/sqrt(1+A^2)) +
    B/sqrt(1+B^2) atan(A/sqrt(1+B^2))], A = B = w/h."""
    A = B = w / h
    fc = (A / math.sqrt(1 + A * A) * math.atan(B / math.sqrt(1 + A * A)) +
          B / math.sqrt(1 + B * B) * math.atan(A / math.sqrt(1 + B * B))) / (2 * math.pi)
    return 4 * fc


def direct_light_scene(albedo=0.6, le=15.0, w=0.5, h=1.0):
    b = _Builder()
    b.quad((-3, 0, -3), (6, 0, 0), (0, 0, 6), (albedo,) * 3)                      # floor (y = 0)
    b.quad((-w, h, -w), (2 * w, 0, 0), (0, 0, 2 * w), (0, 0, 0), (le,) * 3)     # emitter (y = h)
    cam = dict(position=(0.0, 0.3, 0.0), target=(0.0, 0.0, 0.0), up=(0.0, 0.0, -1.0), fov=1.0)
    return b.mesh(cam, "direct")


@pytest.mark.parametrize("integrator,rel", [(1, 0.01), (0, 0.03)])
def test_direct_light_view_factor(integrator, rel):
    albedo, le = 0.6, 15.0
    F = view_factor_point_to_centered_square(0.5, 1.0)
    assert F == pytest.approx(0.23945, abs=1e-4)
    s = O.Scene(direct_light_scene(albedo, le))
    r = s.render(8, 8, 4096, max_depth=2, rr_start=8, integrator=integrator, rng_mode=1, seed=1)
    got = r.mean()[..., :3].mean()
    assert got == pytest.approx(albedo * le * F, rel=rel)


def nogap_cornell():
    m = scenes.cornell_ref()
    li = np.nonzero(m.emission.max(1) > 0)[0]
    v = m.vertices.copy()
    v[np.unique(m.indices[li]), 1] = F32(1.5)
    return Mesh(v, m.indices, m.albedo, m.emission, m.camera, "cornell_lowlight")


def test_nee_and_reference_loop_have_equal_expectation():
    s = O.Scene(nogap_cornell())
    means = {0: [], 1: []}
    for integ in (0, 1):
        for seed in range(6):
            r = s.render(24, 24, 256, max_depth=5, rr_start=3, integrator=integ, rng_mode=1, seed=100 + seed)
            means[integ].append(r.mean()[..., :3].mean())
    a, b = np.array(means[1]), np.array(means[0])
    se = math.sqrt(a.var(ddof=1) / a.size + b.var(ddof=1) / b.size)
    assert abs(a.mean() - b.mean()) < 5 * se + 1e-3 * b.mean()
    assert a.std() < b.std()     # next-event estimation reduces variance here

"""The oracle's adjoint, pinned without the GPU.

- The frozen-decision evaluator (Tape.eval) at the base parameters reproduces
  the oracle renderer's image, so it evaluates the same estimator.
- The FP64 reverse-mode adjoint agrees with central differences of the frozen
  evaluator (eps 1e-3, FP64, Richardson-extrapolated so that the O(eps^2)
  truncation term, about 4e-5 relative for vertex parameters, cancels) on
  random albedo and vertex parameters.
- The albedo gradient satisfies Euler's identity: every NEE term found at
  bounce b carries b + 1 albedo factors, so sum(albedo * dl/dalbedo) equals
  sum_k (k - 1) (l_k - l_{k-1}) with l_k the loss of forward renders of depth
  k (no Russian roulette): forward renders only, no derivative code.
- The reference integrator has no continuous dependence on vertex positions
  (emission is constant per triangle): its vertex gradient is zero.
"""
import numpy as np
import pytest

from oracle import oracle as O
from quasi_amd import scenes

W, H, SPP = 24, 20, 3


def small_scene(stage="init"):
    return scenes.diff_sphere(frequency=4, stage=stage)


def weights(seed=3):
    g = np.random.Generator(np.random.PCG64(seed)).normal(size=(H, W, 4)).astype(np.float32)
    g[..., 3] = 0.0
    return g


def loss_from_render(scene, dimg, **kw):
    r = scene.render(W, H, SPP, precision=64, **kw)
    return float((r.mean("beauty")[..., :3] * dimg[..., :3].astype(np.float64)).sum())


@pytest.mark.parametrize("integrator,max_depth,rr_start", [(1, 4, 3), (1, 6, 1), (0, 5, 2)])
def test_frozen_eval_reproduces_render(integrator, max_depth, rr_start):
    mesh = small_scene()
    sc = O.Scene(mesh)
    dimg = weights()
    kw = dict(max_depth=max_depth, rr_start=rr_start, integrator=integrator, rng_mode=1, seed=7)
    ref = loss_from_render(sc, dimg, **kw)
    tape = sc.tape(W, H, SPP, dimg, **kw)
    assert tape.paths == W * H * SPP
    got = tape.eval()
    assert abs(got - ref) <= 1e-12 * max(1.0, abs(ref)), (got, ref)


def rel_err(a, b, floor):
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


@pytest.mark.parametrize("integrator,max_depth,rr_start", [(1, 4, 3), (1, 6, 2)])
def test_fp64_adjoint_matches_central_differences(integrator, max_depth, rr_start):
    mesh = small_scene()
    sc = O.Scene(mesh)
    dimg = weights(5)
    kw = dict(max_depth=max_depth, rr_start=rr_start, integrator=integrator, rng_mode=1, seed=11)
    ga, gv = sc.adjoint(W, H, SPP, dimg, precision=64, **kw)
    tape = sc.tape(W, H, SPP, dimg, **kw)
    rng = np.random.Generator(np.random.PCG64(6))
    nt = mesh.num_triangles
    sphere_v = 10 * 4 * 4 + 2
    # parameters the image actually depends on, plus random ones
    hit_tris = np.nonzero(np.abs(ga).sum(axis=1) > 0)[0]
    hit_verts = np.nonzero(np.abs(gv[:sphere_v]).sum(axis=1) > 0)[0]
    assert hit_tris.size > 20 and hit_verts.size > 20
    checks = []
    for t in rng.choice(hit_tris, 12, replace=False):
        c = int(rng.integers(3))
        checks.append(("albedo", int(t), c, ga[t, c]))
    for v in rng.choice(hit_verts, 12, replace=False):
        c = int(rng.integers(3))
        checks.append(("vertex", int(v), c, gv[v, c]))
    for t in rng.choice(nt, 4, replace=False):
        checks.append(("albedo", int(t), 0, ga[t, 0]))
    scale = max(np.abs(ga).max(), np.abs(gv).max())
    for kind, i, c, g in checks:
        fd = tape.fd(kind, i, c, eps=1e-3, richardson=True)
        err = abs(g - fd) / max(abs(fd), 1e-3 * scale)
        assert err <= 1e-5, (kind, i, c, g, fd, err)


def test_albedo_gradient_euler_identity():
    mesh = small_scene("target")
    sc = O.Scene(mesh)
    dimg = weights(9)
    D = 5
    kw = dict(rr_start=D, integrator=1, rng_mode=1, seed=3)
    ga, _ = sc.adjoint(W, H, SPP, dimg, max_depth=D, precision=64, **kw)
    alb = mesh.albedo.astype(np.float64)
    lhs = float((alb * ga).sum())
    losses = [loss_from_render(sc, dimg, max_depth=k, **kw) for k in range(1, D + 1)]
    rhs = sum((k - 1) * (losses[k - 1] - losses[k - 2]) for k in range(2, D + 1))
    assert abs(lhs - rhs) <= 1e-9 * max(1.0, abs(rhs)), (lhs, rhs)


def test_reference_integrator_has_no_vertex_gradient():
    mesh = small_scene()
    sc = O.Scene(mesh)
    dimg = weights(2)
    kw = dict(max_depth=5, rr_start=2, integrator=0, rng_mode=1, seed=1)
    ga, gv = sc.adjoint(W, H, SPP, dimg, precision=64, **kw)
    assert np.abs(ga).max() > 0
    assert np.abs(gv).max() == 0.0
    tape = sc.tape(W, H, SPP, dimg, **kw)
    t = int(np.argmax(np.abs(ga).sum(axis=1)))
    for c in range(3):
        fd = tape.fd("albedo", t, c)
        assert abs(ga[t, c] - fd) <= 1e-6 * max(abs(fd), 1e-12)


def test_fp32_adjoint_tracks_fp64():
    mesh = small_scene()
    sc = O.Scene(mesh)
    dimg = weights(4)
    kw = dict(max_depth=4, rr_start=3, integrator=1, rng_mode=1, seed=5)
    ga64, gv64 = sc.adjoint(W, H, SPP, dimg, precision=64, **kw)
    ga32, gv32 = sc.adjoint(W, H, SPP, dimg, precision=32, **kw)
    for a, b in ((ga32, ga64), (gv32, gv64)):
        assert np.abs(a - b).sum() <= 1e-3 * np.abs(b).sum()


def test_adjoint_skips_zero_weight_pixels_and_strides():
    mesh = small_scene()
    sc = O.Scene(mesh)
    dimg = weights(8)
    kw = dict(max_depth=4, rr_start=3, integrator=1, rng_mode=1, seed=5)
    masked = dimg.copy().reshape(-1, 4)
    masked[np.arange(W * H) % 3 != 0] = 0.0
    a1, v1 = sc.adjoint(W, H, SPP, masked.reshape(H, W, 4), precision=64, **kw)
    a2, v2 = sc.adjoint(W, H, SPP, dimg, precision=64, pixel_stride=3, **kw)
    assert np.allclose(a1, a2, rtol=1e-12, atol=1e-15) and np.allclose(v1, v2, rtol=1e-12, atol=1e-15)

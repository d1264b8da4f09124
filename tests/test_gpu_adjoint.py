"""GPU path-replay adjoint (Q_plugin_adjoint) against the oracle.

The gradient is that of l = sum_pixels dL/dI . I over the replayed samples,
with each path's discrete decisions frozen (see the oracle's
oracle_adjoint.hpp, whose FP64 adjoint and frozen-decision evaluator are
pinned on the CPU by tests/test_oracle_adjoint.py).

Bars (written here, checked below):
- against the FP64 CPU adjoint of the same samples: relative L1 error of the
  albedo and vertex gradient vectors <= 1e-3;
- against FP64 central differences (eps 1e-3) of the frozen estimator on 64
  random parameters (seed 6) among those the image depends on: relative error
  <= 1e-3 per parameter, with the denominator floored at 1e-2 times the
  largest |gradient| of the 64 (a parameter whose gradient cancels to near
  zero is judged against the scale of the others).  The differences are
  Richardson-extrapolated ((4 D(eps/2) - D(eps)) / 3): on the 9,680-triangle
  sphere (edges about 0.036) a vertex step of 1e-3 leaves an O(eps^2)
  truncation error of up to 1.5% in the plain central difference, which the
  FP64 CPU adjoint shows as well; it falls as eps^2 (1.5e-4 at eps 1e-4);
- the L2-target mode returns the loss sum (M - T)^2 of the accumulated mean M
  to 1e-5 and the same gradient as the image-gradient mode fed 2 (M - T).
"""
import numpy as np
import pytest

from oracle import oracle as O
from quasi_amd import abi, scenes
from quasi_amd.plugin import HipRenderer, make_camera

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def setup(r, mesh, w, h, spp, depth=4, rr_start=3, integrator=1, seed=1):
    r.set_scene(mesh)
    r.resize(w, h)
    c = mesh.camera
    r.set_camera(c["position"], c["target"], c["up"], c["fov"])
    r.set_params(spp_per_render=spp, max_depth=depth, rr_start=rr_start, rng_mode=1, integrator=integrator,
                 seed=seed, engine=abi.Q_ENGINE_WAVEFRONT, bvh_width=4, block_size=256, queue_chunk=64)
    r.render(1)
    return r.readback_aov()["beauty"]


def rel_l1(a, b):
    return float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))


def weights(w, h, seed):
    g = np.random.Generator(np.random.PCG64(seed)).normal(size=(h, w, 4)).astype(np.float32)
    g[..., 3] = 0.0
    return g


@pytest.mark.parametrize("integrator,depth,rr_start", [(1, 4, 3), (1, 6, 2), (0, 5, 2)])
def test_adjoint_matches_fp64_oracle(integrator, depth, rr_start):
    mesh = scenes.diff_sphere(frequency=8, stage="init")
    w, h, spp = 96, 80, 8
    say(f"[adjoint-vs-oracle] integrator={integrator} depth={depth} rr={rr_start} {w}x{h} spp={spp}")
    dimg = weights(w, h, 21)
    with HipRenderer(w, h) as r:
        setup(r, mesh, w, h, spp, depth, rr_start, integrator, seed=1)
        ga, gv, res = r.adjoint(dimg, spp)
        assert res.paths == w * h * spp and res.kernel_ms > 0
    sc = O.Scene(mesh)
    ra, rv = sc.adjoint(w, h, spp, dimg, max_depth=depth, rr_start=rr_start, integrator=integrator, rng_mode=1,
                        seed=1, precision=64)
    ea, ev = rel_l1(ga, ra), rel_l1(gv, rv)
    say(f"[adjoint-vs-oracle] rel-L1 albedo {ea:.2e} vertices {ev:.2e}")
    assert np.abs(ra).max() > 0
    assert ea <= 1e-3
    if integrator == 0:
        assert np.abs(gv).max() == 0.0 and np.abs(rv).max() == 0.0
    else:
        assert np.abs(rv).max() > 0 and ev <= 1e-3


def test_adjoint_matches_central_differences_cfg4():
    """The differentiable-rendering scene (9,680-triangle icosphere, ground,
    2 lights; initial albedo 0.5 and jittered vertices) at 128 x 128, 8 spp,
    depth 4, with dL/dI = 2 (I - T) for the target scene's image T."""
    w = h = 128
    spp = 8
    init = scenes.diff_sphere(22, "init")
    target = scenes.diff_sphere(22, "target")
    with HipRenderer(w, h) as r:
        T = setup(r, target, w, h, spp)
        M = setup(r, init, w, h, spp)
        dimg = np.zeros((h, w, 4), dtype=np.float32)
        dimg[..., :3] = 2.0 * (M[..., :3] - T[..., :3])
        ga, gv, _ = r.adjoint(dimg, spp)
    sc = O.Scene(init)
    kw = dict(max_depth=4, rr_start=3, integrator=1, rng_mode=1, seed=1)
    ra, rv = sc.adjoint(w, h, spp, dimg, precision=64, **kw)
    say(f"[adjoint-fd] rel-L1 vs FP64 adjoint: albedo {rel_l1(ga, ra):.2e} vertices {rel_l1(gv, rv):.2e}")
    assert rel_l1(ga, ra) <= 1e-3 and rel_l1(gv, rv) <= 1e-3
    tape = sc.tape(w, h, spp, dimg, **kw)
    nsph_t, nsph_v = 9680, 4842
    rng = np.random.Generator(np.random.PCG64(6))
    live_t = np.nonzero(np.abs(ra[:nsph_t]).sum(axis=1) > 0)[0]
    live_v = np.nonzero(np.abs(rv[:nsph_v]).sum(axis=1) > 0)[0]
    params = [("albedo", int(t), int(rng.integers(3))) for t in rng.choice(live_t, 32, replace=False)]
    params += [("vertex", int(v), int(rng.integers(3))) for v in rng.choice(live_v, 32, replace=False)]
    gpu = np.array([ga[i, c] if k == "albedo" else gv[i, c] for k, i, c in params])
    fd = np.array([tape.fd(k, i, c, eps=1e-3, richardson=True) for k, i, c in params])
    floor = 1e-2 * np.abs(fd).max()
    err = np.abs(gpu - fd) / np.maximum(np.abs(fd), floor)
    worst = int(np.argmax(err))
    say(f"[adjoint-fd] 64 params: max rel err {err.max():.2e} (param {params[worst]}, gpu {gpu[worst]:.6g}, "
        f"fd {fd[worst]:.6g}); median {np.median(err):.2e}")
    assert err.max() <= 1e-3
    cpu = np.array([ra[i, c] if k == "albedo" else rv[i, c] for k, i, c in params])
    cerr = np.abs(gpu - cpu) / np.maximum(np.abs(cpu), 1e-2 * np.abs(cpu).max())
    say(f"[adjoint-fd] same 64 params vs the FP64 CPU adjoint: max rel err {cerr.max():.2e}")
    assert cerr.max() <= 1e-3


def test_l2_target_mode_and_full_size_subsample():
    """Full differentiable-rendering workload (512 x 512, 32 spp, depth 4):
    the L2-target mode's loss and gradient, and the gradient of every 13th
    pixel against the FP64 CPU adjoint."""
    w = h = 512
    spp = 32
    init = scenes.diff_sphere(22, "init")
    target = scenes.diff_sphere(22, "target")
    say("[adjoint-l2] 512x512 spp=32")
    with HipRenderer(w, h) as r:
        T = setup(r, target, w, h, spp)
        M = setup(r, init, w, h, spp)
        ga_l2, gv_l2, res = r.adjoint(T, spp, mode=abi.Q_ADJOINT_L2_TARGET)
        loss = float(((M[..., :3].astype(np.float64) - T[..., :3]) ** 2).sum())
        say(f"[adjoint-l2] loss {res.loss:.6f} (host {loss:.6f}), adjoint {res.kernel_ms:.2f} ms")
        assert abs(res.loss - loss) <= 1e-5 * loss
        dimg = np.zeros((h, w, 4), dtype=np.float32)
        dimg[..., :3] = 2.0 * (M[..., :3] - T[..., :3])
        ga, gv, _ = r.adjoint(dimg, spp)
        assert rel_l1(ga_l2, ga) <= 1e-4 and rel_l1(gv_l2, gv) <= 1e-4
        sub = dimg.reshape(-1, 4).copy()
        sub[np.arange(w * h) % 13 != 0] = 0.0
        sub = sub.reshape(h, w, 4)
        gs, vs, _ = r.adjoint(sub, spp)
    sc = O.Scene(init)
    ra, rv = sc.adjoint(w, h, spp, sub, max_depth=4, rr_start=3, integrator=1, rng_mode=1, seed=1, precision=64)
    say(f"[adjoint-l2] every 13th pixel: rel-L1 albedo {rel_l1(gs, ra):.2e} vertices {rel_l1(vs, rv):.2e}")
    assert rel_l1(gs, ra) <= 1e-3 and rel_l1(vs, rv) <= 1e-3


def test_adjoint_full_size_cfg4_fd_and_fp64():
    """The differentiable-rendering workload at its stated size (512 x 512,
    32 spp, depth 4, dL/dI = 2 (I - T)): the whole image's gradient against
    the FP64 CPU adjoint, and 64 random parameters (seed 6) against FP64
    central differences of the frozen estimator (Richardson-extrapolated),
    at eps 1e-3 where that difference has converged (it agrees with eps 1e-4
    to 1e-4) and at eps 1e-4 for all 64."""
    w = h = 512
    spp = 32
    init = scenes.diff_sphere(22, "init")
    target = scenes.diff_sphere(22, "target")
    say("[adjoint-full] 512x512 spp=32 depth 4")
    with HipRenderer(w, h) as r:
        T = setup(r, target, w, h, spp)
        M = setup(r, init, w, h, spp)
        dimg = np.zeros((h, w, 4), dtype=np.float32)
        dimg[..., :3] = 2.0 * (M[..., :3] - T[..., :3])
        ga, gv, res = r.adjoint(dimg, spp)
        say(f"[adjoint-full] GPU adjoint {res.kernel_ms:.2f} ms of kernels")
    sc = O.Scene(init)
    kw = dict(max_depth=4, rr_start=3, integrator=1, rng_mode=1, seed=1)
    ra, rv = sc.adjoint(w, h, spp, dimg, precision=64, **kw)
    ea, ev = rel_l1(ga, ra), rel_l1(gv, rv)
    say(f"[adjoint-full] whole image vs FP64 adjoint: rel-L1 albedo {ea:.2e} vertices {ev:.2e}")
    assert ea <= 1e-3 and ev <= 1e-3
    tape = sc.tape(w, h, spp, dimg, **kw)
    nsph_t, nsph_v = 9680, 4842
    rng = np.random.Generator(np.random.PCG64(6))
    live_t = np.nonzero(np.abs(ra[:nsph_t]).sum(axis=1) > 0)[0]
    live_v = np.nonzero(np.abs(rv[:nsph_v]).sum(axis=1) > 0)[0]
    params = [("albedo", int(t), int(rng.integers(3))) for t in rng.choice(live_t, 32, replace=False)]
    params += [("vertex", int(v), int(rng.integers(3))) for v in rng.choice(live_v, 32, replace=False)]
    gpu = np.array([ga[i, c] if k == "albedo" else gv[i, c] for k, i, c in params])
    # eps 1e-3 (Richardson-extrapolated) and, to see whether that difference
    # has converged, eps 1e-4: at full size a few vertices near the jittered
    # sphere's silhouette have gradients of 100+ that are far from linear over
    # 1e-3 (vertex 77: 131.12 at eps 1e-3, 133.887 at eps 1e-4, where the FP64
    # and the FP32 CPU adjoints also are)
    fd3 = np.array([tape.fd(k, i, c, eps=1e-3, richardson=True) for k, i, c in params])
    fd4 = np.array([tape.fd(k, i, c, eps=1e-4, richardson=True) for k, i, c in params])
    tape.close()
    floor = 1e-2 * np.abs(fd4).max()
    err = np.abs(gpu - fd4) / np.maximum(np.abs(fd4), floor)
    conv = np.abs(fd3 - fd4) / np.maximum(np.abs(fd4), floor) <= 1e-4
    err3 = np.abs(gpu - fd3) / np.maximum(np.abs(fd3), floor)
    worst = int(np.argmax(err))
    say(f"[adjoint-full] 64 params vs FD (eps 1e-4): max rel err {err.max():.2e} (param {params[worst]}, "
        f"gpu {gpu[worst]:.6g}, fd {fd4[worst]:.6g}); median {np.median(err):.2e}")
    say(f"[adjoint-full] eps 1e-3 converged on {int(conv.sum())} of 64 params; there the max rel err is "
        f"{err3[conv].max():.2e}; unconverged: "
        f"{[(params[i], round(float(fd3[i]), 4), round(float(fd4[i]), 4)) for i in np.nonzero(~conv)[0]]}")
    assert err.max() <= 1e-3
    assert conv.sum() >= 48 and err3[conv].max() <= 1e-3


def test_adjoint_deep_paths_use_the_largest_lds_layout():
    """Depth 12 (the adjoint kernel's 16-segment LDS layout, with the
    visibility flag packed into the triangle word) against the FP64 oracle."""
    mesh = scenes.diff_sphere(frequency=8, stage="init")
    w, h, spp = 48, 40, 4
    say(f"[adjoint-deep] depth 12 rr 2 {w}x{h} spp={spp}")
    dimg = weights(w, h, 31)
    with HipRenderer(w, h) as r:
        setup(r, mesh, w, h, spp, depth=12, rr_start=2, integrator=1, seed=3)
        ga, gv, _ = r.adjoint(dimg, spp)
    sc = O.Scene(mesh)
    ra, rv = sc.adjoint(w, h, spp, dimg, max_depth=12, rr_start=2, integrator=1, rng_mode=1, seed=3, precision=64)
    ea, ev = rel_l1(ga, ra), rel_l1(gv, rv)
    say(f"[adjoint-deep] rel-L1 albedo {ea:.2e} vertices {ev:.2e}")
    assert np.abs(ra).max() > 0 and np.abs(rv).max() > 0
    assert ea <= 1e-3 and ev <= 1e-3


def tilted_ground_sphere():
    """The init-stage sphere of diff_sphere (frequency 8) on a ground quad
    tilted about both horizontal axes: the ground's two triangles are far
    above the split threshold and not axis-aligned, so the builders cut each
    into four references (bvh_split.hpp), whose records share the triangle's
    id; the oracle traces its own unsplit BVH."""
    v, idx = scenes.icosphere(8)
    v = (v + np.random.Generator(np.random.PCG64(5)).normal(0.0, 0.01, size=v.shape)).astype(np.float32)
    nt = idx.shape[0]
    b = scenes._Builder()
    b.v, b.i = [v], [idx]
    b.a, b.e = [np.full((nt, 3), 0.5, np.float32)], [np.zeros((nt, 3), np.float32)]
    b.quad((-2.0, -0.3, 2.0), (4.0, 0.4, 0.0), (0.0, 0.2, -4.0), (0.6, 0.6, 0.6))
    b.quad((-0.9, 2.0, -0.1), (0.6, 0.0, 0.0), (0.0, 0.0, 0.6), (0.0, 0.0, 0.0), (12.0, 12.0, 12.0))
    b.quad((0.6, 1.6, -0.6), (0.4, 0.0, 0.0), (0.0, 0.0, 0.4), (0.0, 0.0, 0.0), (10.0, 8.0, 6.0))
    return b.mesh(scenes.DIFF_CAMERA, "tilted_ground_sphere")


def test_adjoint_over_split_references_matches_fp64_oracle():
    """Gradients are keyed by the input triangle (the record's id), not by the
    BVH reference: with the ground split into references, the albedo and
    vertex gradients still match the FP64 oracle, ground vertices included."""
    mesh = tilted_ground_sphere()
    w, h, spp = 96, 80, 8
    say(f"[adjoint-split] tilted ground, depth 4 rr 3 {w}x{h} spp={spp}")
    dimg = weights(w, h, 41)
    with HipRenderer(w, h) as r:
        setup(r, mesh, w, h, spp, depth=4, rr_start=3, integrator=1, seed=2)
        st = r.stats()
        say(f"[adjoint-split] references {st.bvh_references} for {st.num_triangles} triangles")
        assert st.bvh_references >= st.num_triangles + 4   # both ground triangles in 3 or more pieces
        ga, gv, _ = r.adjoint(dimg, spp)
    sc = O.Scene(mesh)
    ra, rv = sc.adjoint(w, h, spp, dimg, max_depth=4, rr_start=3, integrator=1, rng_mode=1, seed=2, precision=64)
    ea, ev = rel_l1(ga, ra), rel_l1(gv, rv)
    ground = slice(len(mesh.vertices) - 12, len(mesh.vertices) - 8)   # the ground quad's four vertices
    eg = rel_l1(gv[ground], rv[ground])
    say(f"[adjoint-split] rel-L1 albedo {ea:.2e} vertices {ev:.2e} ground vertices {eg:.2e}")
    assert np.abs(rv[ground]).max() > 0
    assert ea <= 1e-3 and ev <= 1e-3 and eg <= 1e-3


def test_adjoint_row_bands_and_sample_blocks_equal_one_pass():
    """Passes bounded by max_pass_paths: blocks of 3 samples, and single
    samples of 7-row bands (the last band 3 rows), give the one-pass gradient
    to FP32 atomic-order rounding, and both match the FP64 oracle."""
    mesh = scenes.diff_sphere(frequency=8, stage="init")
    w, h, spp = 96, 80, 8
    dimg = weights(w, h, 41)
    say(f"[adjoint-bands] {w}x{h} spp={spp}")
    with HipRenderer(w, h) as r:
        setup(r, mesh, w, h, spp, seed=2)
        ga, gv, res = r.adjoint(dimg, spp)
        assert (res.passes, res.bands) == (1, 1)
        gb, vb, rb = r.adjoint(dimg, spp, max_pass_paths=w * h * 3)
        assert (rb.passes, rb.bands) == (3, 1)
        gc, vc, rc = r.adjoint(dimg, spp, max_pass_paths=w * 7)
        assert (rc.passes, rc.bands) == (spp * 12, 12)
        assert rc.working_bytes < res.working_bytes
        say(f"[adjoint-bands] working bytes: one pass {res.working_bytes}, 7-row bands {rc.working_bytes}")
    for a, b in ((gb, ga), (gc, ga)):
        assert rel_l1(a, b) <= 1e-6
    for a, b in ((vb, gv), (vc, gv)):
        assert rel_l1(a, b) <= 1e-6
    sc = O.Scene(mesh)
    ra, rv = sc.adjoint(w, h, spp, dimg, max_depth=4, rr_start=3, integrator=1, rng_mode=1, seed=2, precision=64)
    assert rel_l1(gc, ra) <= 1e-3 and rel_l1(vc, rv) <= 1e-3


def test_adjoint_albedo_parameter_map():
    """Triangles sharing albedo parameters: the device's FP64 reduction onto
    the map equals the per-triangle gradients summed on the host; unmapped
    triangles contribute nothing."""
    mesh = scenes.diff_sphere(frequency=8, stage="init")
    w, h, spp = 64, 48, 4
    nt = mesh.num_triangles
    pmap = (np.arange(nt, dtype=np.uint32) // 5).astype(np.uint32)
    pmap[::7] = abi.Q_ADJOINT_NO_PARAM
    npar = int(pmap[pmap != abi.Q_ADJOINT_NO_PARAM].max()) + 1
    dimg = weights(w, h, 43)
    with HipRenderer(w, h) as r:
        setup(r, mesh, w, h, spp, seed=4)
        ga, gv, _ = r.adjoint(dimg, spp)
        gp, gvp, _ = r.adjoint(dimg, spp, albedo_param=pmap, num_albedo_params=npar)
        with pytest.raises(RuntimeError):
            bad = pmap.copy()
            bad[3] = npar
            r.adjoint(dimg, spp, albedo_param=bad, num_albedo_params=npar)
    keep = pmap != abi.Q_ADJOINT_NO_PARAM
    host = np.stack([np.bincount(pmap[keep], weights=ga[keep, c], minlength=npar) for c in range(3)], axis=1)
    say(f"[adjoint-map] {npar} parameters: rel-L1 vs host sums {rel_l1(gp, host):.2e}")
    assert gp.shape == (npar, 3)
    assert rel_l1(gp, host) <= 1e-6
    assert rel_l1(gvp, gv) <= 1e-6


def test_adjoint_seed_and_sample_range():
    """Samples 4..7 replayed with an explicit seed equal the gradient of an
    oracle render of those samples with that seed."""
    mesh = scenes.diff_sphere(frequency=6, stage="target")
    w, h = 64, 48
    dimg = weights(w, h, 5)
    with HipRenderer(w, h) as r:
        setup(r, mesh, w, h, 2, seed=0)
        ga, gv, _ = r.adjoint(dimg, 4, sample_begin=4, seed=77)
    sc = O.Scene(mesh)
    ra, rv = sc.adjoint(w, h, 4, dimg, sample_offset=4, max_depth=4, rr_start=3, integrator=1, rng_mode=1, seed=77,
                        precision=64)
    assert rel_l1(ga, ra) <= 1e-3 and rel_l1(gv, rv) <= 1e-3


def test_adjoint_errors():
    with HipRenderer(32, 32) as r:
        r.set_scene(scenes.diff_sphere(frequency=2))
        img = np.zeros((32, 32, 4), dtype=np.float32)
        with pytest.raises(RuntimeError):
            r.adjoint(img, 1)                       # no frame rendered yet (no camera)
        r.render(1)
        with pytest.raises(RuntimeError):
            r.adjoint(img, 0)                       # no samples
        ga, gv, res = r.adjoint(img, 1)             # zero image gradient: zero gradients
        assert not ga.any() and not gv.any()

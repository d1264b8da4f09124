"""GPU edge cases against the FP64 oracle, through the C ABI: ragged and tiny
images, a one-triangle scene without lights, a camera that sees nothing, a
scene of no triangles, depth 1, and a sample count that does not divide into whole passes.  Same bars
as test_gpu_parity.py (relative L1 <= 1e-4 on the same samples, ids >= 99.99%
with only near-tie disagreements, exact camera-ray counts)."""
import numpy as np
import pytest

from oracle import oracle as O
from quasi_amd import abi, scenes
from quasi_amd.plugin import HipRenderer, make_camera
from tests.parity import check_ids, rel_l1

pytestmark = pytest.mark.gpu

REL_BAR = 1e-4
ENGINES = [abi.Q_ENGINE_MEGAKERNEL, abi.Q_ENGINE_WAVEFRONT, abi.Q_ENGINE_PERSISTENT]
ENGINE_IDS = ["megakernel", "wavefront", "single-launch"]


def say(*a):
    print(*a, flush=True)


def cam_of(mesh):
    c = mesh.camera
    return make_camera(c["position"], c["target"], c["up"], c["fov"])


def render_and_compare(mesh, w, h, spp, engine, label, **kw):
    params = dict(max_depth=kw.pop("max_depth", 4), rr_start=kw.pop("rr_start", 3), rng_mode=1,
                  integrator=kw.pop("integrator", 1), seed=kw.pop("seed", 3))
    with HipRenderer(w, h) as r:
        r.set_scene(mesh)
        r.camera = cam_of(mesh)
        r.set_params(spp_per_render=spp, flags=abi.Q_FLAG_RECORD_IDS, id_sample=0, engine=engine, **params, **kw)
        r.render(1)
        aov = r.readback_aov()
        ids = r.primary_ids()
        st = r.stats()
    s = O.Scene(mesh)
    ref = s.render(w, h, spp, precision=64, **params)
    errs = {k: rel_l1(aov[k], ref.mean(k)) for k in ("beauty", "albedo", "normal", "depth")}
    frac, nbad, not_ties = check_ids(ids, ref.ids, mesh, s,
                                     dict(width=w, height=h, rng_mode=1, seed=params["seed"], camera=mesh.camera))
    say(f"[{label}] {w}x{h} spp={spp} rel-L1 {errs} ids {frac:.6f} rays {st.rays_camera}/{st.rays_extension}/"
        f"{st.rays_shadow} oracle {ref.counts.camera_rays}/{ref.counts.extension_rays}/{ref.counts.shadow_rays}")
    for k, e in errs.items():
        assert e <= REL_BAR, (label, k, e)
    assert frac >= 0.9999 and not not_ties, (label, frac, not_ties[:3])
    assert st.rays_camera == ref.counts.camera_rays == w * h * spp
    return aov, st, ref


@pytest.mark.parametrize("engine", ENGINES, ids=ENGINE_IDS)
@pytest.mark.parametrize("w,h", [(1, 1), (7, 5), (33, 17)])
def test_tiny_and_ragged_images(w, h, engine):
    render_and_compare(scenes.cornell36(2), w, h, 5, engine, "ragged")


@pytest.mark.parametrize("engine", ENGINES, ids=ENGINE_IDS)
@pytest.mark.parametrize("integrator", [0, 1])
def test_one_triangle_without_lights(engine, integrator):
    """NEE with an empty light table and the reference loop on a scene of one
    triangle: nothing emits, so every image is black, but the hits, AOVs and
    ray counts still follow the oracle."""
    m = scenes.Mesh(vertices=np.array([[-1, 0, -1], [1, 0, -1], [0, 1.5, -1]], dtype=np.float32),
                    indices=np.array([[0, 1, 2]], dtype=np.uint32),
                    albedo=np.array([[0.7, 0.5, 0.3]], dtype=np.float32),
                    emission=np.zeros((1, 3), dtype=np.float32),
                    camera=dict(position=(0.0, 0.6, 2.0), target=(0.0, 0.6, -1.0), up=(0.0, 1.0, 0.0), fov=40.0),
                    name="one")
    aov, st, ref = render_and_compare(m, 24, 16, 3, engine, "one-triangle", integrator=integrator)
    assert float(np.abs(aov["beauty"][..., :3]).max()) == 0.0
    assert (aov["depth"][..., 0] > 0).any()


@pytest.mark.parametrize("engine", ENGINES, ids=ENGINE_IDS)
def test_camera_that_sees_nothing(engine):
    m = scenes.cornell36(1)
    m.camera = dict(position=(0.0, 1.0, 3.5), target=(0.0, 1.0, 10.0), up=(0.0, 1.0, 0.0), fov=40.0)
    aov, st, ref = render_and_compare(m, 20, 12, 2, engine, "miss")
    assert st.rays_extension == 0 and st.rays_shadow == 0
    assert float(np.abs(aov["beauty"][..., :3]).max()) == 0.0


def empty_mesh():
    m = scenes.cornell36(1)
    z3 = np.zeros((0, 3), dtype=np.float32)
    return scenes.Mesh(z3, np.zeros((0, 3), dtype=np.uint32), z3, z3, m.camera, "empty")


@pytest.mark.parametrize("engine", ENGINES, ids=ENGINE_IDS)
@pytest.mark.parametrize("integrator", [0, 1], ids=["reference", "nee"])
def test_empty_scene(engine, integrator):
    """A scene of no triangles: every camera ray misses, nothing else is
    traced, the image and the first-hit records are zero and every id is a
    miss."""
    w, h, spp = 20, 12, 3
    say(f"[empty] engine {engine} integrator {integrator}")
    with HipRenderer(w, h) as r:
        r.set_scene(empty_mesh())
        r.camera = cam_of(empty_mesh())
        r.set_params(spp_per_render=spp, flags=abi.Q_FLAG_RECORD_IDS, id_sample=0, engine=engine, max_depth=4,
                     rr_start=3, rng_mode=1, integrator=integrator, seed=3)
        r.render(2)
        aov = r.readback_aov()
        ids = r.primary_ids()
        st = r.stats()
    assert st.num_triangles == 0 and st.bvh_nodes == 0
    assert st.rays_camera == 2 * w * h * spp and st.rays_extension == 0 and st.rays_shadow == 0
    for k in ("beauty", "albedo", "normal", "depth"):
        assert float(np.abs(aov[k][..., :3]).max()) == 0.0, k
    assert (ids == 0xFFFFFFFF).all()


@pytest.mark.parametrize("engine", ENGINES, ids=ENGINE_IDS)
def test_depth_one(engine):
    aov, st, ref = render_and_compare(scenes.cornell36(3), 32, 24, 4, engine, "depth1", max_depth=1)
    assert st.rays_extension == 0 and st.rays_shadow == 0


def test_ragged_sample_passes():
    """7 samples in passes of 3 (3 + 3 + 1) on the wavefront engine."""
    render_and_compare(scenes.cornell36(4), 40, 30, 7, abi.Q_ENGINE_WAVEFRONT, "passes", samples_per_pass=3,
                       max_depth=5)


def test_rejected_scene_keeps_the_previous_one():
    """A scene with a vertex index out of range is rejected before anything
    on the device changes: renders and the adjoint then still use the previous
    scene (device arrays and host mesh alike).  Renders agree bit for bit; the
    adjoint's FP32 atomic sums depend on the order waves arrive in, so two
    calls agree to a few units in the last place (rtol 1e-6)."""
    good = scenes.cornell36(1)
    bad = scenes.Mesh(good.vertices.copy(), good.indices.copy(), good.albedo.copy(), good.emission.copy(),
                      dict(good.camera), "bad")
    bad.indices[5, 1] = good.vertices.shape[0]
    w, h = 32, 24
    target = np.full((h, w, 4), 0.25, dtype=np.float32)
    with HipRenderer(w, h) as r:
        r.set_scene(good)
        r.camera = cam_of(good)
        r.set_params(spp_per_render=2, max_depth=4, rr_start=3, rng_mode=1, integrator=1, seed=4)
        r.render(2)
        before = r.readback()
        ga0, gv0, _ = r.adjoint(target, 2, mode=abi.Q_ADJOINT_L2_TARGET)
        with pytest.raises(RuntimeError):
            r.set_scene(bad)
        say(f"[rejected] {r.messages[-1]}")
        r.set_params(seed=4)                     # restart the accumulation
        r.render(2)
        after = r.readback()
        ga1, gv1, _ = r.adjoint(target, 2, mode=abi.Q_ADJOINT_L2_TARGET)
    assert np.array_equal(before, after)
    assert np.allclose(ga0, ga1, rtol=1e-6, atol=1e-9) and np.allclose(gv0, gv1, rtol=1e-6, atol=1e-9)
    assert np.abs(ga0).max() > 0 and np.abs(gv0).max() > 0


def test_rejected_large_scene_through_the_gpu_build():
    """A 20k-triangle mesh goes to the GPU build: its corner gather finds the
    index out of range, the host path then rejects the scene with its own
    check, and the previous scene keeps rendering bit for bit."""
    good = scenes.soup(20_000, seed=5)
    bad = scenes.Mesh(good.vertices.copy(), good.indices.copy(), good.albedo.copy(), good.emission.copy(),
                      dict(good.camera), "bad")
    bad.indices[12_345, 2] = 10 * good.vertices.shape[0]
    w, h = 40, 30
    with HipRenderer(w, h) as r:
        r.set_scene(good)
        assert r.stats().bvh_builder == abi.Q_BVH_BUILD_GPU
        r.camera = cam_of(good)
        r.set_params(spp_per_render=2, max_depth=4, rr_start=3, rng_mode=1, integrator=1, seed=6)
        r.render(1)
        before = r.readback()
        with pytest.raises(RuntimeError):
            r.set_scene(bad)
        log = "\n".join(r.messages[-3:])
        say(f"[rejected-gpu] {log}")
        assert "GPU build failed (vertex index out of range)" in log and "out of range" in r.messages[-1]
        r.set_params(seed=6)
        r.render(1)
        after = r.readback()
    assert np.array_equal(before, after)


def test_scene_beyond_the_quantised_range_traces_the_full_nodes():
    """One triangle at x = 1e21 beside the room: the root's extent needs a
    quantisation step above 2^60, so the plugin says so and traces the
    128-byte nodes under the default parameters; the image equals the
    128-byte traversal's bit for bit."""
    room = scenes.cornell36(3)
    nv = room.vertices.shape[0]
    far = np.array([[1e21, 0.0, 0.0], [1e21, 1.0, 0.0], [1e21, 0.0, 1.0]], dtype=np.float32)
    mesh = scenes.Mesh(np.concatenate([room.vertices, far]),
                       np.concatenate([room.indices, np.array([[nv, nv + 1, nv + 2]], dtype=np.uint32)]),
                       np.concatenate([room.albedo, np.full((1, 3), 0.5, np.float32)]),
                       np.concatenate([room.emission, np.zeros((1, 3), np.float32)]), dict(room.camera), "far")
    c = mesh.camera
    imgs = []
    logs = []
    for chunk in (0, 1):
        with HipRenderer(40, 32, log=logs.append) as r:
            r.set_scene(mesh)
            r.camera = make_camera(c["position"], c["target"], c["up"], c["fov"])
            r.set_params(spp_per_render=4, max_depth=5, rr_start=3, rng_mode=1, integrator=1, seed=3,
                         engine=abi.Q_ENGINE_WAVEFRONT, queue_chunk=chunk)
            r.render(1)
            imgs.append(r.readback())
    print("[quant-range]", [m for m in logs if "quantised" in m], flush=True)
    assert any("beyond the quantised nodes' range" in m for m in logs)
    assert np.array_equal(imgs[0], imgs[1])
    assert float(imgs[0][..., :3].sum()) > 0.0

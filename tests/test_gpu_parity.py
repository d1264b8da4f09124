"""GPU parity: libquasi_hip.so (through the C ABI) against the FP64 oracle on
the same samples.

Bars (written here, checked below):
- beauty and AOV images: relative L1 error <= 1e-4 (see tests/parity.py);
- primary-hit triangle ids: >= 99.99% agreement, every disagreement a near tie;
- ray counts: camera rays exact, extension/shadow rays within 0.1% (paths may
  diverge after a near-tie or a rounding-sensitive decision).
"""
import numpy as np
import pytest

from oracle import oracle as O
from quasi_amd import abi, scenes
from quasi_amd.plugin import HipRenderer, make_camera
from tests.parity import check_ids, rel_l1

pytestmark = pytest.mark.gpu

REL_BAR = 1e-4


def say(*a):
    print(*a, flush=True)


def oracle_render(mesh, w, h, spp, **kw):
    s = O.Scene(mesh)
    return s, s.render(w, h, spp, precision=64, **kw)


def compare(r, ref, mesh, oscene, params, label, check_aovs=True):
    aov = r.readback_aov()
    ids = r.primary_ids()
    errs = {"beauty": rel_l1(aov["beauty"], ref.mean("beauty"))}
    if check_aovs:
        for k in ("albedo", "normal", "depth"):
            errs[k] = rel_l1(aov[k], ref.mean(k))
    frac, nbad, not_ties = check_ids(ids, ref.ids, mesh, oscene, params)
    say(f"[{label}] rel-L1 {errs} ids agree {frac:.6f} ({nbad} differ, {len(not_ties)} not near ties)")
    for k, e in errs.items():
        assert e <= REL_BAR, (label, k, e)
    assert np.allclose(aov["beauty"][..., 3], 1.0)
    assert frac >= 0.9999, (label, frac)
    assert not not_ties, (label, not_ties[:5])
    return aov, ids


def cam_of(mesh):
    c = mesh.camera
    return make_camera(c["position"], c["target"], c["up"], c["fov"])


ENGINES = [abi.Q_ENGINE_MEGAKERNEL, abi.Q_ENGINE_WAVEFRONT, abi.Q_ENGINE_PERSISTENT]
ENGINE_IDS = ["megakernel", "wavefront", "single-launch"]


@pytest.mark.parametrize("engine", ENGINES, ids=ENGINE_IDS)
def test_reference_mode_default_scene(engine):
    """The plugin's own default scene and defaults (plugin.mm:556-592): 1 spp
    per Q_plugin_render, reference integrator and seeding, depth 5."""
    w, h, frames = 128, 128, 16
    say(f"[ref-default] engine={engine} {w}x{h} frames={frames}")
    with HipRenderer(w, h) as r:
        p = r.get_params()
        assert (p.spp_per_render, p.max_depth, p.rr_start, p.rng_mode, p.integrator) == (1, 5, 3, 0, 0)
        r.set_params(flags=abi.Q_FLAG_RECORD_IDS, id_sample=0, engine=engine)
        r.render(frames)
        mesh = scenes.cornell_ref()
        params = dict(width=w, height=h, rng_mode=0, camera=mesh.camera)
        oscene, ref = oracle_render(mesh, w, h, frames, max_depth=5, rr_start=3, rng_mode=0, integrator=0)
        compare(r, ref, mesh, oscene, params, "ref-default")
        st = r.stats()
        assert st.sample_count == frames and st.num_triangles == 32 and st.num_lights == 2
        c = ref.counts
        assert st.rays_camera == c.camera_rays == w * h * frames
        assert abs(st.rays_extension - c.extension_rays) <= 1e-3 * c.extension_rays
        assert st.rays_shadow == 0


@pytest.mark.parametrize("engine", ENGINES, ids=ENGINE_IDS)
def test_nee_cornell36_config(engine):
    """36-triangle Cornell-style scene, 256x256, 16 spp, depth 4, NEE, hashed seeding."""
    w = h = 256
    mesh = scenes.cornell36(0)
    say(f"[nee-cornell36] engine={engine} {w}x{h} spp=16")
    with HipRenderer(w, h) as r:
        r.set_scene(mesh)
        r.camera = cam_of(mesh)
        r.set_params(spp_per_render=16, max_depth=4, rr_start=3, rng_mode=1, integrator=1, seed=0,
                     flags=abi.Q_FLAG_RECORD_IDS, id_sample=0, engine=engine)
        r.render(1)
        params = dict(width=w, height=h, rng_mode=1, seed=0, camera=mesh.camera)
        oscene, ref = oracle_render(mesh, w, h, 16, max_depth=4, rr_start=3, rng_mode=1, integrator=1, seed=0)
        compare(r, ref, mesh, oscene, params, "nee-cornell36")
        st = r.stats()
        c = ref.counts
        assert st.rays_camera == c.camera_rays
        assert abs(st.rays_extension - c.extension_rays) <= 1e-3 * c.extension_rays
        assert abs(st.rays_shadow - c.shadow_rays) <= 1e-3 * c.shadow_rays


@pytest.mark.parametrize("engine", ENGINES, ids=ENGINE_IDS)
@pytest.mark.parametrize("ntri,w,h,spp", [(200_000, 160, 90, 4)])
def test_nee_soup_bvh(ntri, w, h, spp, engine):
    """Triangle soup (the 1M-triangle workload's generator at reduced size),
    depth 8 with Russian roulette from bounce 3: exercises the BVH."""
    mesh = scenes.soup(ntri, seed=1)
    say(f"[nee-soup] engine={engine} ntri={ntri} {w}x{h} spp={spp}")
    with HipRenderer(w, h) as r:
        r.set_scene(mesh)
        r.camera = cam_of(mesh)
        r.set_params(spp_per_render=spp, max_depth=8, rr_start=3, rng_mode=1, integrator=1, seed=0,
                     flags=abi.Q_FLAG_RECORD_IDS, id_sample=0, engine=engine)
        r.render(1)
        params = dict(width=w, height=h, rng_mode=1, seed=0, camera=mesh.camera)
        oscene, ref = oracle_render(mesh, w, h, spp, max_depth=8, rr_start=3, rng_mode=1, integrator=1, seed=0)
        compare(r, ref, mesh, oscene, params, "nee-soup")
        st = r.stats()
        say(f"[nee-soup] bvh nodes={st.bvh_nodes} depth={st.bvh_depth} build_ms={st.bvh_build_ms:.1f}")
        c = ref.counts
        assert st.rays_camera == c.camera_rays
        assert abs(st.rays_extension - c.extension_rays) <= 1e-3 * c.extension_rays
        assert abs(st.rays_shadow - c.shadow_rays) <= 1e-3 * c.shadow_rays


def sliver_mesh(n, seed):
    """Long thin triangles across the soup's cube (each far larger than the
    average leaf cell, so the builders split nearly every one into four
    references) under the soup's four lights."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1.0, 1.0, size=(n, 3))
    b = a + rng.uniform(-0.6, 0.6, size=(n, 3))
    c = b + rng.normal(scale=0.01, size=(n, 3))
    tri = np.stack([a, b, c], axis=1).astype(np.float32)
    bld = scenes._Builder()
    bld.triangles(tri, rng.uniform(0.0, 1.0, size=(n, 3)).astype(np.float32), np.zeros((n, 3), np.float32))
    for cx, cy, cz in scenes.SOUP_LIGHT_CENTERS:
        bld.quad((cx - 0.25, cy, cz - 0.25), (0.5, 0, 0), (0, 0, 0.5), (0.0, 0.0, 0.0), scenes.LIGHT_E)
    return bld.mesh(scenes.SOUP_CAMERA, f"slivers{n}")


def test_split_references_match_oracle_on_slivers():
    """A scene whose every sliver is split into references (bvh_split.hpp):
    both builders give the same references and the same images, bit for bit,
    and the image matches the oracle's (which traces its own unsplit BVH).
    The oracle comparison is at depth 2 (camera rays, one bounce and its
    shadow rays): these crossing slivers make deeper paths diverge between
    FP32 and FP64 arithmetic (the oracle's own FP32 and FP64 renders differ by
    7e-6 at depth 2, 8e-5 at depth 3 and 2.4e-4 at depth 8, against 5e-7 at
    depth 8 on the soup), so depth 8 is compared between the builders only."""
    mesh = sliver_mesh(20_000, 7)
    w, h, spp = 128, 72, 4
    params = dict(width=w, height=h, rng_mode=1, seed=0, camera=mesh.camera)
    oscene, ref = oracle_render(mesh, w, h, spp, max_depth=2, rr_start=3, rng_mode=1, integrator=1, seed=0)
    out = {}
    for builder in (abi.Q_BVH_BUILD_HOST, abi.Q_BVH_BUILD_GPU):
        say(f"[slivers] builder={builder} {mesh.num_triangles} triangles {w}x{h} spp={spp}")
        with HipRenderer(w, h) as r:
            r.set_params(bvh_builder=builder)
            r.set_scene(mesh)
            r.camera = cam_of(mesh)
            common = dict(spp_per_render=spp, rr_start=3, rng_mode=1, integrator=1, seed=0,
                          flags=abi.Q_FLAG_RECORD_IDS, id_sample=0, engine=abi.Q_ENGINE_WAVEFRONT,
                          bvh_builder=builder)
            r.set_params(max_depth=2, **common)
            r.render(1)
            st = r.stats()
            assert st.bvh_builder == builder
            say(f"[slivers] references {st.bvh_references} for {st.num_triangles} triangles, "
                f"nodes {st.bvh_nodes}")
            assert st.num_triangles == mesh.num_triangles
            # the slivers split (four references each at most), the 8 light triangles stay whole
            assert 3 * 20_000 < st.bvh_references <= 4 * 20_000 + 8
            compare(r, ref, mesh, oscene, params, f"slivers-{builder}")
            r.set_params(max_depth=8, **common)
            r.render(1)
            out[builder] = (st.bvh_references, r.readback_aov()["beauty"].copy(), r.primary_ids().copy())
    (n_h, img_h, ids_h), (n_g, img_g, ids_g) = out[abi.Q_BVH_BUILD_HOST], out[abi.Q_BVH_BUILD_GPU]
    assert n_h == n_g
    assert np.array_equal(img_h, img_g) and np.array_equal(ids_h, ids_g)


def test_render_is_deterministic_and_passes_compose():
    """Two identical renders are bitwise equal; 8 calls of 1 spp match one call
    of 8 spp (same sample indices) to FP32 summation-order rounding."""
    w = h = 64
    mesh = scenes.cornell36(0)
    say("[determinism]")
    with HipRenderer(w, h) as r:
        r.set_scene(mesh)
        r.camera = cam_of(mesh)
        r.set_params(spp_per_render=8, max_depth=4, rng_mode=1, integrator=1)
        r.render(1)
        a = r.readback()
        r.set_params(spp_per_render=8, max_depth=4, rng_mode=1, integrator=1)
        r.render(1)
        b = r.readback()
        assert np.array_equal(a, b)
        r.set_params(spp_per_render=1, max_depth=4, rng_mode=1, integrator=1)
        r.render(8)
        c = r.readback()
        assert np.allclose(a, c, rtol=2e-6, atol=1e-6)


def test_camera_dirty_resets_and_resize():
    """camera_dirty restarts the accumulation (plugin.mm:643-646); a frame of
    another size reallocates it (plugin.mm:638-640)."""
    say("[camera-dirty/resize]")
    with HipRenderer(48, 40) as r:
        assert r.readback() is None               # nothing rendered yet
        r.render(3)
        r.set_camera((0.3, 1.1, 3.2), (0.0, 1.0, 0.0))
        r.render(2)
        moved = r.readback()
        assert r.stats().sample_count == 2
        with HipRenderer(48, 40) as fresh:
            fresh.set_camera((0.3, 1.1, 3.2), (0.0, 1.0, 0.0))
            fresh.render(2)
            assert np.array_equal(moved, fresh.readback())
        r.resize(32, 24)
        r.render(1)
        img = r.readback()
        assert img.shape == (24, 32, 4)
        assert r.stats().sample_count == 1


@pytest.mark.parametrize("integrator", [0, 1])
def test_wavefront_matches_megakernel(integrator):
    """Same samples through both engines: FP32 images within 1e-6 relative L1,
    identical primary ids and ray counts (except rare rounding-order divergence)."""
    mesh = scenes.soup(100_000, seed=3)
    w, h, spp = 128, 72, 6
    say(f"[wave-vs-mega] integrator={integrator}")
    out = {}
    for engine in ENGINES[:2]:
        with HipRenderer(w, h) as r:
            r.set_scene(mesh)
            r.camera = cam_of(mesh)
            r.set_params(spp_per_render=spp, max_depth=8, rr_start=3, rng_mode=1, integrator=integrator, seed=2,
                         flags=abi.Q_FLAG_RECORD_IDS, id_sample=3, engine=engine, samples_per_pass=2)
            r.render(1)
            out[engine] = (r.readback_aov(), r.primary_ids(), r.stats())
    (am, im, sm), (aw, iw, sw) = out[ENGINES[0]], out[ENGINES[1]]
    for k in ("beauty", "albedo", "normal", "depth"):
        assert rel_l1(aw[k], am[k]) <= 1e-6, k
    assert (im == iw).mean() >= 0.9999
    assert sm.rays_camera == sw.rays_camera == w * h * spp
    assert abs(int(sm.rays_extension) - int(sw.rays_extension)) <= 1e-4 * sm.rays_extension
    assert abs(int(sm.rays_shadow) - int(sw.rays_shadow)) <= 1e-4 * max(1, sm.rays_shadow)


@pytest.mark.parametrize("integrator,width,block", [(1, 4, 256), (1, 2, 128), (0, 4, 128)])
def test_single_launch_equals_wavefront(integrator, width, block):
    """Q_ENGINE_PERSISTENT traces every path as the wavefront kernels do and
    accumulates in the same order: every AOV, the primary ids and the ray
    counts equal the wavefront engine's bit for bit."""
    mesh = scenes.soup(150_000, seed=8)
    w, h, spp = 160, 96, 8
    say(f"[single-launch] integrator={integrator} width={width} block={block}")
    out = {}
    for engine in (abi.Q_ENGINE_WAVEFRONT, abi.Q_ENGINE_PERSISTENT):
        with HipRenderer(w, h) as r:
            r.set_scene(mesh)
            r.camera = cam_of(mesh)
            r.set_params(spp_per_render=spp, max_depth=8, rr_start=3, rng_mode=1, integrator=integrator, seed=5,
                         flags=abi.Q_FLAG_RECORD_IDS, id_sample=2, engine=engine, bvh_width=width, block_size=block,
                         samples_per_pass=4)
            r.render(1)
            out[engine] = (r.readback_aov(), r.primary_ids(), r.stats())
    (aw, iw, sw), (ap, ip, sp) = out[abi.Q_ENGINE_WAVEFRONT], out[abi.Q_ENGINE_PERSISTENT]
    for k in ("beauty", "albedo", "normal", "depth"):
        assert np.array_equal(aw[k], ap[k]), k
    assert np.array_equal(iw, ip)
    assert (sw.rays_camera, sw.rays_extension, sw.rays_shadow) == (sp.rays_camera, sp.rays_extension, sp.rays_shadow)
    names = {sp.kernels[k].name.decode(): sp.kernels[k].launches for k in range(sp.num_kernels)}
    assert names["paths_pass"] == 2 and names["wf_extend"] == 0


def test_wavefront_pass_split_and_determinism():
    """The wavefront result does not depend on how the samples are split into
    passes (up to FP32 summation order) and repeats bit for bit."""
    mesh = scenes.cornell36(0)
    w, h = 96, 64
    say("[wave-passes]")
    imgs = []
    with HipRenderer(w, h) as r:
        r.set_scene(mesh)
        r.camera = cam_of(mesh)
        for per_pass in (1, 3, 8, 8):
            r.set_params(spp_per_render=8, max_depth=4, rng_mode=1, integrator=1, engine=abi.Q_ENGINE_WAVEFRONT,
                         samples_per_pass=per_pass)
            r.render(1)
            imgs.append(r.readback())
    assert np.array_equal(imgs[2], imgs[3])
    for im in imgs[:2]:
        assert np.allclose(im, imgs[2], rtol=2e-6, atol=1e-6)


TRACE_CONFIGS = ([(w, b, c) for w in (2, 4) for b in (64, 128, 256) for c in (1, 64, 256)] +
                 [(w, 128, abi.Q_QUEUE_FLAT) for w in (2, 4)] +
                 [(4, b, c) for c in (abi.Q_QUEUE_LDS_TOP, abi.Q_QUEUE_QUANT) for b in (128, 256)])


def test_all_trace_configs_render_identically():
    """Every compiled traversal variant (BVH width 2/4, block 64/128/256, queue
    chunk: per-lane fetch, 64, 256, the non-persistent form, per-lane fetch
    with the top BVH4 levels staged in LDS, and that over the quantised
    64-byte nodes) yields the same
    image bit for bit: the closest hit is unique under the tie rule and
    occlusion does not depend on which blocker is found."""
    mesh = scenes.soup(150_000, seed=8)
    w, h = 96, 64
    say("[trace-configs]")
    ref = None
    with HipRenderer(w, h) as r:
        r.set_scene(mesh)
        r.camera = cam_of(mesh)
        for (bw, blk, ch) in TRACE_CONFIGS:
            r.set_params(spp_per_render=4, max_depth=6, rr_start=3, rng_mode=1, integrator=1,
                         engine=abi.Q_ENGINE_WAVEFRONT, bvh_width=bw, block_size=blk, queue_chunk=ch)
            r.render(1)
            img = r.readback()
            if ref is None:
                ref = img
            assert np.array_equal(img, ref), (bw, blk, ch)


@pytest.mark.parametrize("width", [2, 4])
def test_bvh4_parity_with_oracle(width):
    mesh = scenes.soup(200_000, seed=1)
    w, h, spp = 120, 80, 4
    say(f"[bvh-width] width={width}")
    with HipRenderer(w, h) as r:
        r.set_scene(mesh)
        r.camera = cam_of(mesh)
        r.set_params(spp_per_render=spp, max_depth=8, rr_start=3, rng_mode=1, integrator=1, seed=5,
                     flags=abi.Q_FLAG_RECORD_IDS, id_sample=1, engine=abi.Q_ENGINE_WAVEFRONT, bvh_width=width)
        r.render(1)
        params = dict(width=w, height=h, rng_mode=1, seed=5, camera=mesh.camera)
        oscene = O.Scene(mesh)
        ref = oscene.render(w, h, spp, max_depth=8, rr_start=3, rng_mode=1, integrator=1, seed=5, id_sample=1)
        aov = r.readback_aov()
        assert rel_l1(aov["beauty"], ref.mean()) <= REL_BAR
        frac, nbad, not_ties = check_ids(r.primary_ids(), ref.ids, mesh, oscene, params, sample=1)
        assert frac >= 0.9999 and not not_ties


def test_unsupported_trace_config_is_rejected():
    with HipRenderer(16, 16) as r:
        with pytest.raises(RuntimeError):
            r.set_params(bvh_width=3)
        with pytest.raises(RuntimeError):
            r.set_params(block_size=96)
        with pytest.raises(RuntimeError):
            r.set_params(queue_chunk=100)
        r.set_params(bvh_width=4, block_size=128, queue_chunk=256)


def test_params_and_stats_accept_older_struct_sizes():
    """A caller built against the header before the bvh_builder fields passes
    the older struct sizes; the plugin reads and writes only those bytes."""
    import ctypes as C
    old_params = abi.Q_render_params.bvh_builder.offset
    old_stats = abi.Q_render_stats.bvh_builder.offset
    with HipRenderer(16, 16) as r:
        p = abi.Q_render_params()
        p.struct_size = old_params
        p.bvh_builder = 77                      # beyond the old size: must stay untouched
        assert r._get_params(r.plugin.handle, C.byref(p)) == abi.Q_EXT_OK
        assert p.struct_size == old_params and p.bvh_builder == 77 and p.max_depth == 5
        p.max_depth = 3
        assert r._set_params(r.plugin.handle, C.byref(p)) == abi.Q_EXT_OK   # bvh_builder read as 0 (auto)
        q = r.get_params()
        assert q.max_depth == 3 and q.bvh_builder == 0
        s = abi.Q_render_stats()
        s.struct_size = old_stats
        s.bvh_sah_cost = -1.0
        assert r._get_stats(r.plugin.handle, C.byref(s)) == abi.Q_EXT_OK
        assert s.bvh_sah_cost == -1.0 and s.num_triangles == 32
        p.struct_size = old_params - 4
        assert r._get_params(r.plugin.handle, C.byref(p)) == abi.Q_EXT_ERR_ARGUMENT

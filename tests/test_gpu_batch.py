"""Batched scenes (Q_plugin_set_batch / render_batch): one pass covers every
scene with shared queues and is replayed from a captured hipGraph.

Bars: every scene of a batch equals, bit for bit, a single-scene wavefront
render of it with the same seed and samples per pass (same kernels, same
per-path arithmetic, same accumulation order); graph replays interleaved with
eager work each match the eager result; a batch scene matches the FP64 oracle
to the image bar (relative L1 <= 1e-4).
"""
import numpy as np
import pytest

from oracle import oracle as O
from quasi_amd import abi, scenes
from quasi_amd.plugin import HipRenderer, make_camera
from tests.parity import check_ids, rel_l1

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def small_batch():
    out = scenes.batch_scenes(num_scenes=6, base_seed=77, min_tris=40, max_tris=20000,
                              resolutions=(24, 40, 64))
    # one non-square scene and one without lights (all soup) for coverage
    m, w, h, s = out[1]
    out[1] = (m, 48, 32, s)
    soup = scenes.soup(3000, seed=9)
    soup.emission[:] = 0
    out.append((soup, 32, 24, 4242))
    return out


PARAMS = dict(max_depth=6, rr_start=3, rng_mode=1, integrator=1, engine=abi.Q_ENGINE_WAVEFRONT, samples_per_pass=1,
              bvh_width=4, block_size=128, queue_chunk=64, flags=abi.Q_FLAG_RECORD_IDS, id_sample=1)


def plain_render(r, mesh, w, h, seed, spp, params=None):
    r.set_scene(mesh)
    r.resize(w, h)
    c = mesh.camera
    r.camera = make_camera(c["position"], c["target"], c["up"], c["fov"])
    r.set_params(spp_per_render=spp, seed=seed, **(params or PARAMS))
    r.render(1)
    return r.readback_aov(), r.primary_ids()


@pytest.mark.parametrize("builder", [abi.Q_BVH_BUILD_AUTO, abi.Q_BVH_BUILD_GPU], ids=["auto", "gpu"])
def test_batch_equals_plain_per_scene(builder):
    """With the GPU builder the batch's scenes are built on the device and
    their arrays concatenated there; with the automatic one these small scenes
    are built on host threads."""
    batch = small_batch()
    spp = 4
    say(f"[batch==plain] {len(batch)} scenes, builder {builder}")
    with HipRenderer(16, 16) as r:
        r.set_params(spp_per_render=1, seed=0, bvh_builder=builder, **PARAMS)
        r.set_batch(batch)
        r.render_batch(spp)
        got = [(r.readback_batch(i), r.batch_ids(i)) for i in range(len(batch))]
        for i, (mesh, w, h, seed) in enumerate(batch):
            aov, ids = plain_render(r, mesh, w, h, seed, spp)
            for k in ("beauty", "albedo", "normal", "depth"):
                assert np.array_equal(got[i][0][k], aov[k]), (i, k)
            assert np.array_equal(got[i][1], ids), i


def test_single_launch_batch_equals_wavefront_batch():
    """A batch rendered with one persistent launch per pass (Q_ENGINE_PERSISTENT,
    captured in the pass graph) equals the wavefront launch sequence's batch
    bit for bit, scene by scene, and each scene its single-scene render."""
    batch = small_batch()
    spp = 4
    say(f"[batch single-launch] {len(batch)} scenes")
    got = {}
    for engine in (abi.Q_ENGINE_WAVEFRONT, abi.Q_ENGINE_PERSISTENT):
        with HipRenderer(16, 16) as r:
            r.set_params(spp_per_render=1, seed=0, **dict(PARAMS, engine=engine, block_size=128))
            r.set_batch(batch)
            r.render_batch(spp)
            got[engine] = [(r.readback_batch(i), r.batch_ids(i)) for i in range(len(batch))]
            st = r.stats()
            say(f"[batch single-launch] engine {engine}: rays {st.rays_camera}/{st.rays_extension}/{st.rays_shadow}")
    for i in range(len(batch)):
        (aw, iw), (ap, ip) = got[abi.Q_ENGINE_WAVEFRONT][i], got[abi.Q_ENGINE_PERSISTENT][i]
        for k in ("beauty", "albedo", "normal", "depth"):
            assert np.array_equal(aw[k], ap[k]), (i, k)
        assert np.array_equal(iw, ip), i


def test_batch_with_an_empty_scene():
    """A scene of no triangles among others: its image is zero and its ids are
    misses, and the other scenes still equal their plain renders."""
    z3 = np.zeros((0, 3), dtype=np.float32)
    room = scenes.cornell36(5)
    empty = scenes.Mesh(z3, np.zeros((0, 3), dtype=np.uint32), z3, z3, room.camera, "empty")
    batch = [(room, 24, 16, 11), (empty, 16, 16, 12), (scenes.cornell36(6), 16, 24, 13)]
    spp = 3
    say("[batch-empty]")
    with HipRenderer(16, 16) as r:
        r.set_params(spp_per_render=1, seed=0, bvh_builder=abi.Q_BVH_BUILD_GPU, **PARAMS)
        r.set_batch(batch)
        r.render_batch(spp)
        got = [(r.readback_batch(i), r.batch_ids(i)) for i in range(len(batch))]
        for k in ("beauty", "albedo", "normal", "depth"):
            assert float(np.abs(got[1][0][k][..., :3]).max()) == 0.0, k
        assert (got[1][1] == 0xFFFFFFFF).all()
        for i in (0, 2):
            mesh, w, h, seed = batch[i]
            aov, ids = plain_render(r, mesh, w, h, seed, spp)
            for k in ("beauty", "albedo", "normal", "depth"):
                assert np.array_equal(got[i][0][k], aov[k]), (i, k)
            assert np.array_equal(got[i][1], ids), i


def subset(mesh, tris, name):
    """The triangles `tris` of mesh (all vertices kept)."""
    t = np.asarray(tris)
    return scenes.Mesh(mesh.vertices, mesh.indices[t], mesh.albedo[t], mesh.emission[t], dict(mesh.camera), name)


@pytest.mark.parametrize("block", [128, 256])
def test_default_traversal_batch_with_tiny_scenes(block):
    """The default traversal of a batch (quantised nodes; every wave stages the
    top three levels of one scene in LDS) over scenes of one, two, five and
    sixty triangles (with the soup's lights), none, and five thousand: each
    equals its plain render bit for bit.  A one-triangle scene's root has a
    single used slot; the small trees have fewer internal nodes than the 21
    staged slots."""
    soup = scenes.soup(5000, seed=21)
    n = soup.num_triangles
    lights = list(range(n - 8, n))
    z3 = np.zeros((0, 3), dtype=np.float32)
    empty = scenes.Mesh(z3, np.zeros((0, 3), dtype=np.uint32), z3, z3, dict(soup.camera), "empty")
    batch = [(subset(soup, [0], "one"), 16, 16, 1), (empty, 16, 16, 2), (subset(soup, [0, 1], "two"), 24, 16, 3),
             (soup, 48, 32, 4), (subset(soup, list(range(5)) + lights, "five"), 16, 24, 5),
             (subset(soup, list(range(60)) + lights, "sixty"), 32, 32, 6)]
    params = dict(PARAMS, block_size=block, queue_chunk=0)
    spp = 4
    say(f"[batch-tiny] block {block}")
    with HipRenderer(16, 16) as r:
        r.set_params(spp_per_render=1, seed=0, **params)
        r.set_batch(batch)
        r.render_batch(spp)
        got = [(r.readback_batch(i), r.batch_ids(i)) for i in range(len(batch))]
        for i, (mesh, w, h, seed) in enumerate(batch):
            aov, ids = plain_render(r, mesh, w, h, seed, spp, params)
            for k in ("beauty", "albedo", "normal", "depth"):
                assert np.array_equal(got[i][0][k], aov[k]), (i, k)
            assert np.array_equal(got[i][1], ids), i


def test_graph_replays_match_eager_with_work_between():
    batch = small_batch()[:4]
    say("[graph-replay]")
    with HipRenderer(16, 16) as r:
        r.set_params(spp_per_render=1, seed=0, **PARAMS)
        r.set_batch(batch)
        for k in range(1, 4):
            r.render_batch(1)           # one graph replay
            r.synchronize()
            # eager work on the same plugin between replays
            r.render(1)
            r.synchronize()
            imgs = [r.readback_batch(i)["beauty"] for i in range(len(batch))]
            with HipRenderer(16, 16) as e:
                for i, (mesh, w, h, seed) in enumerate(batch):
                    aov, _ = plain_render(e, mesh, w, h, seed, k)
                    assert np.array_equal(imgs[i], aov["beauty"]), (k, i)


def test_batch_scene_matches_oracle():
    batch = small_batch()
    spp = 8
    i = 3
    mesh, w, h, seed = batch[i]
    say(f"[batch-oracle] scene {i} {w}x{h} ntri={mesh.num_triangles}")
    with HipRenderer(16, 16) as r:
        r.set_params(spp_per_render=1, seed=0, **PARAMS)
        r.set_batch(batch)
        r.render_batch(spp)
        got = r.readback_batch(i)
        ids = r.batch_ids(i)
    oscene = O.Scene(mesh)
    ref = oscene.render(w, h, spp, max_depth=6, rr_start=3, rng_mode=1, integrator=1, seed=seed, id_sample=1)
    for k in ("beauty", "albedo", "normal", "depth"):
        assert rel_l1(got[k], ref.mean(k)) <= 1e-4, k
    frac, nbad, not_ties = check_ids(ids, ref.ids, mesh, oscene,
                                     dict(width=w, height=h, rng_mode=1, seed=seed, camera=mesh.camera), sample=1)
    assert frac >= 0.9999 and not not_ties


def test_batch_stats_and_errors():
    batch = small_batch()[:3]
    with HipRenderer(16, 16) as r:
        r.set_params(spp_per_render=1, seed=0, **PARAMS)
        with pytest.raises(RuntimeError):
            r.render_batch(1)                       # no batch yet
        r.set_batch(batch)
        r.reset_stats()
        r.render_batch(2)
        st = r.stats()
        total_pix = sum(w * h for _, w, h, _ in batch)
        assert st.rays_camera == 2 * total_pix
        kern = {st.kernels[k].name.decode(): st.kernels[k] for k in range(st.num_kernels)}
        assert kern["batch_pass"].launches == 2
        with pytest.raises(RuntimeError):
            r.readback_batch(len(batch))            # out of range
        r.set_params(spp_per_render=1, seed=0, **dict(PARAMS, samples_per_pass=2))
        with pytest.raises(RuntimeError):
            r.render_batch(3)                       # not a multiple of 2 samples per pass
        r.render_batch(4)


def test_autotune_batch(tmp_path, monkeypatch):
    monkeypatch.setenv("QUASI_HIP_TUNE_CACHE", str(tmp_path / "tune.jsonl"))
    batch = small_batch()[:4]
    say("[autotune]")
    with HipRenderer(16, 16) as r:
        r.set_params(spp_per_render=1, seed=0, **dict(PARAMS, flags=abi.Q_FLAG_RECORD_IDS | abi.Q_FLAG_TIME_KERNELS))
        r.set_batch(batch)
        r.render_batch(1)
        before = r.stats()
        res = r.autotune(abi.Q_AUTOTUNE_BATCH, probe_spp=1)
        # every traversal shape once; a 1-sample probe leaves no choice of pass size
        assert res.num_entries == 24 and res.best < res.num_entries and not res.from_cache
        assert all(res.entries[k].samples_per_pass == 1 for k in range(res.num_entries))
        # the probes leave the caller's statistics as they were
        after = r.stats()
        assert before.rays_camera > 0
        assert (after.rays_camera, after.rays_extension, after.rays_shadow) == \
            (before.rays_camera, before.rays_extension, before.rays_shadow)
        assert [(after.kernels[k].launches, after.kernels[k].total_ms) for k in range(after.num_kernels)] == \
            [(before.kernels[k].launches, before.kernels[k].total_ms) for k in range(before.num_kernels)]
        best = res.entries[res.best]
        assert all(best.ms <= res.entries[k].ms for k in range(res.num_entries))
        p = r.get_params()
        assert (p.bvh_width, p.block_size, p.queue_chunk) == (best.bvh_width, best.block_size, best.queue_chunk)
        assert (tmp_path / "tune.jsonl").exists()
        again = r.autotune(abi.Q_AUTOTUNE_BATCH, probe_spp=1, flags=abi.Q_AUTOTUNE_USE_CACHE)
        assert again.from_cache == 1 and again.entries[0].bvh_width == best.bvh_width
        # the tuned configuration still renders the same image as any other
        r.render_batch(2)
        img = r.readback_batch(0)["beauty"]
    with HipRenderer(16, 16) as e:
        mesh, w, h, seed = batch[0]
        aov, _ = plain_render(e, mesh, w, h, seed, 2)
        assert np.array_equal(img, aov["beauty"])

"""Q_plugin_render_noalloc against Q_plugin_render (the entry point it mirrors).

Both render the same scene, camera, samples and samples per pass with the
wavefront engine; Q_plugin_render's readback is the accumulated sum divided by
the sample count in FP32, so the no-alloc sums divided the same way must equal
it bit for bit (bit-exact: the same kernels add the same values in the same
order).  Two calls in a row, the second on a smaller image; sizes beyond the
fixed working area are rejected before anything is launched.
"""
import numpy as np
import pytest

from quasi_amd import abi, scenes
from quasi_amd.plugin import HipRenderer, make_camera
from tests.hip_mem import DeviceArray

pytestmark = pytest.mark.gpu

PARAMS = dict(max_depth=6, rr_start=3, rng_mode=1, integrator=1, seed=3, engine=abi.Q_ENGINE_WAVEFRONT,
              bvh_width=4, block_size=256, queue_chunk=1)


def say(*a):
    print(*a, flush=True)


def plain(r, w, h, spp, per_pass):
    r.resize(w, h)
    r.set_params(spp_per_render=spp, samples_per_pass=per_pass, **PARAMS)
    r.render(1)
    return r.readback_aov()


def noalloc(r, w, h, spp):
    npix = w * h
    with DeviceArray(4 * npix * 16) as buf:
        rc = r.render_noalloc(buf.ptr, w, h, spp, sample_begin=0)
        assert rc == abi.Q_EXT_OK, (rc, r.messages[-1:])
        r.synchronize()
        sums = buf.to_numpy(np.float32).reshape(4, h, w, 4)
    return sums


def test_noalloc_equals_render_on_two_calls():
    mesh = scenes.room_with_soup(30000, seed=5)
    c = mesh.camera
    with HipRenderer(64, 64) as r:
        r.set_scene(mesh)
        r.camera = make_camera(c["position"], c["target"], c["up"], c["fov"])
        for (w, h, spp, per_pass) in ((200, 150, 8, 4), (96, 64, 3, 2)):
            say(f"[noalloc] {w}x{h} spp={spp} samples_per_pass={per_pass}")
            ref = plain(r, w, h, spp, per_pass)
            sums = noalloc(r, w, h, spp)
            for k, name in enumerate(("beauty", "albedo", "normal", "depth")):
                got = sums[k] / np.float32(spp)
                assert np.array_equal(got, ref[name]), (w, h, name, float(np.abs(got - ref[name]).max()))
            assert np.all(sums[0][..., 3] == spp)
            say(f"[noalloc] {w}x{h}: all four AOVs equal Q_plugin_render's bit for bit")


def test_noalloc_rejects_passes_beyond_the_working_area():
    # the buffers are large enough for the requests, so that a render the
    # checks failed to refuse would still stay in bounds
    with HipRenderer(32, 32) as r:
        r.set_params(samples_per_pass=2, **PARAMS)
        # 1920 x 1080 x 2 paths per pass exceed the ~2.8M paths of the area
        with DeviceArray(4 * 1920 * 1080 * 16) as buf:
            assert r.render_noalloc(buf.ptr, 1920, 1080, 2) == abi.Q_EXT_ERR_ARGUMENT
        r.set_params(samples_per_pass=0, **PARAMS)
        with DeviceArray(4 * 4096 * 4096 * 16) as buf:
            assert r.render_noalloc(buf.ptr, 4096, 4096, 1) == abi.Q_EXT_ERR_ARGUMENT
        assert r.render_noalloc(0, 64, 64, 1) == abi.Q_EXT_ERR_ARGUMENT
        assert any("working area" in m for m in r.messages)

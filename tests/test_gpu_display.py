"""The display path (the reference's tonemap_frag, plugin.mm:334-350, and
the host's orbit camera, main.cpp:39-112):
Q_plugin_readback_display against numpy on the plugin's own float mean, the
host's PPM against the same export, and the orbit camera with its per-frame
camera_dirty reset (plugin.mm:643-646) against a one-frame render at the final
camera.  The tonemap's powf is the device library's, so a value may round to
the neighbouring 8-bit code where numpy's powf differs in the last place."""
import ctypes as C
import subprocess

import numpy as np
import pytest

from quasi_amd import abi, scenes
from quasi_amd.plugin import HipRenderer, make_camera
from tests.conftest import ROOT

pytestmark = pytest.mark.gpu

HOST = ROOT / "quasi_amd" / "bin" / "quasi_host"


def say(*a):
    print(*a, flush=True)


def tonemap_numpy(mean):
    m = mean[..., :3].astype(np.float32)
    c = m / (m + np.float32(1.0))
    g = np.power(c, np.float32(1.0) / np.float32(2.2))
    return np.rint(np.clip(g, 0, 1) * np.float32(255.0)).astype(np.uint8)


def read_ppm(path):
    data = path.read_bytes()
    head = data.split(b"\n", 3)
    assert head[0] == b"P6" and head[2] == b"255"
    w, h = map(int, head[1].split())
    return np.frombuffer(head[3], dtype=np.uint8).reshape(h, w, 3)


def run_host(args):
    say("[host]", " ".join(args))
    r = subprocess.run([str(HOST), *args], capture_output=True, text=True, timeout=120)
    say(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0
    return r.stdout


@pytest.mark.parametrize("engine", [abi.Q_ENGINE_MEGAKERNEL, abi.Q_ENGINE_WAVEFRONT])
def test_display_is_the_tonemapped_mean(engine):
    mesh = scenes.cornell36(2)
    w, h = 80, 60
    with HipRenderer(w, h) as r:
        r.set_scene(mesh)
        c = mesh.camera
        r.camera = make_camera(c["position"], c["target"], c["up"], c["fov"])
        r.set_params(spp_per_render=4, max_depth=4, rr_start=3, integrator=1, rng_mode=1, seed=2, engine=engine)
        with pytest.raises(RuntimeError):
            r.readback_display()                 # nothing rendered yet: Q_EXT_ERR_STATE
        r.render(3)
        disp = r.readback_display()
        mean = r.readback()
        small = np.empty(4 * w * h - 1, dtype=np.uint8)
        assert r._readback_display(r.plugin.handle, small.ctypes.data_as(C.POINTER(C.c_uint8)),
                                   small.size) == abi.Q_EXT_ERR_ARGUMENT
    ref = tonemap_numpy(mean)
    diff = np.abs(disp[..., :3].astype(int) - ref.astype(int))
    exact = float((diff == 0).mean())
    say(f"[display] engine {engine}: exact {exact:.5f}, max diff {int(diff.max())}, "
        f"mean code {float(disp[..., :3].mean()):.1f}")
    assert (disp[..., 3] == 255).all()
    assert diff.max() <= 1 and exact >= 0.999
    assert disp[..., :3].max() > 0                # the light and lit walls are in view


def test_host_ppm_is_the_display_export(tmp_path):
    w, h, frames = 64, 48, 3
    ppm = tmp_path / "d.ppm"
    run_host(["--render", str(frames), "--width", str(w), "--height", str(h), "--ppm", str(ppm),
              "--out", str(tmp_path / "d.exr")])
    with HipRenderer(w, h) as r:
        r.render(frames)
        disp = r.readback_display()
    assert np.array_equal(read_ppm(ppm), disp[..., :3])


def test_orbit_camera_restarts_accumulation(tmp_path):
    """Three frames turning 0.25 rad of azimuth before the second and third:
    each turn marks the camera dirty, so what is left is one sample at
    azimuth 0.5 -- bit for bit the image of one frame rendered there.  An
    orbit camera at rest (azimuth 0, elevation 0, distance 3.5) is the
    reference host's default camera."""
    common = ["--width", "48", "--height", "40", "--compression", "none"]
    turned, direct = tmp_path / "turned", tmp_path / "direct"
    log = run_host(common + ["--render", "3", "--orbit-step", "0.25", "--out", f"{turned}.exr",
                             "--ppm", f"{turned}.ppm"])
    assert "Saving EXR with AOVs (1 samples)" in log
    run_host(common + ["--render", "1", "--azimuth", "0.5", "--out", f"{direct}.exr", "--ppm", f"{direct}.ppm"])
    assert (tmp_path / "turned.exr").read_bytes() == (tmp_path / "direct.exr").read_bytes()
    assert (tmp_path / "turned.ppm").read_bytes() == (tmp_path / "direct.ppm").read_bytes()
    rest, default = tmp_path / "rest.exr", tmp_path / "default.exr"
    run_host(common + ["--render", "2", "--azimuth", "0", "--out", str(rest)])
    run_host(common + ["--render", "2", "--out", str(default)])
    assert rest.read_bytes() == default.read_bytes()
    assert (tmp_path / "turned.exr").read_bytes() != default.read_bytes()

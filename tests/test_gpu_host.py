"""The headless C++ host (quasi_amd/bin/quasi_host: HIP context, C++ plugin
loader, EXR writer) against the Python harness on the same plugin: the EXR it
writes holds the binary16 rounding of the plugin's float readback, bit for
bit, and depth.Z the FP32 values themselves."""
import subprocess

import numpy as np
import pytest

from quasi_amd import abi, scenes
from quasi_amd.plugin import HipRenderer, make_camera
from tests.conftest import ROOT
from tests.exr_reader import read_exr

pytestmark = pytest.mark.gpu

HOST = ROOT / "quasi_amd" / "bin" / "quasi_host"


def run_host(args):
    print("[host]", " ".join(args), flush=True)
    r = subprocess.run([str(HOST), *args], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    return r.stdout


def check_exr(path, aov):
    _, ch = read_exr(path)
    for name, key, k in (("R", "beauty", 0), ("G", "beauty", 1), ("B", "beauty", 2), ("A", "beauty", 3),
                         ("albedo.R", "albedo", 0), ("albedo.G", "albedo", 1), ("albedo.B", "albedo", 2),
                         ("normal.X", "normal", 0), ("normal.Y", "normal", 1), ("normal.Z", "normal", 2)):
        assert np.array_equal(ch[name].view(np.uint16), aov[key][..., k].astype(np.float16).view(np.uint16)), name
    assert np.array_equal(ch["depth.Z"], aov["depth"][..., 0])


def test_host_default_scene_matches_plugin(tmp_path):
    w, h, frames = 64, 48, 8
    out = tmp_path / "ref.exr"
    log = run_host(["--render", str(frames), "--width", str(w), "--height", str(h), "--out", str(out)])
    assert "Saved:" in log and "HIP Path Tracer" in log
    with HipRenderer(w, h) as r:
        r.render(frames)
        aov = r.readback_aov()
    check_exr(out, aov)


@pytest.mark.parametrize("engine", ["megakernel", "wavefront"])
def test_host_scene_file_nee(tmp_path, engine):
    w = h = 96
    mesh = scenes.cornell36(0)
    sf = tmp_path / "c36.qscene"
    scenes.save_qscene(mesh, sf)
    out = tmp_path / "c36.exr"
    run_host([str(ROOT / "quasi_amd" / "libquasi_hip.so"), "--render", "2", "--width", str(w), "--height", str(h),
              "--scene", str(sf), "--spp-per-render", "4", "--depth", "4", "--integrator", "nee", "--rng", "hashed",
              "--seed", "3", "--engine", engine, "--compression", "none", "--out", str(out)])
    with HipRenderer(w, h) as r:
        r.set_scene(mesh)
        c = mesh.camera
        r.camera = make_camera(c["position"], c["target"], c["up"], c["fov"])
        r.set_params(spp_per_render=4, max_depth=4, rr_start=3, integrator=1, rng_mode=1, seed=3,
                     engine=abi.Q_ENGINE_WAVEFRONT if engine == "wavefront" else abi.Q_ENGINE_MEGAKERNEL)
        r.render(2)
        aov = r.readback_aov()
    check_exr(out, aov)


@pytest.mark.parametrize("engine", ["megakernel", "wavefront"])
def test_host_hot_reload_restarts_identically(tmp_path, engine):
    """--reload-at K: before frame K the host waits for the GPU, the plugin is
    destroyed and its library closed, a fresh copy of the library is opened
    (hot reload, quasi_amd/host/manager.hpp) and handed the scene and
    parameters again, and a new accumulation starts.  After N frames the image
    is that of the N - K frames since the reload: bit for bit the EXR of a run
    of N - K frames without a reload."""
    w = h = 64
    mesh = scenes.cornell36(1)
    sf = tmp_path / "c36.qscene"
    scenes.save_qscene(mesh, sf)
    common = [str(ROOT / "quasi_amd" / "libquasi_hip.so"), "--width", str(w), "--height", str(h), "--scene", str(sf),
              "--spp-per-render", "2", "--depth", "4", "--integrator", "nee", "--rng", "hashed", "--seed", "5",
              "--engine", engine, "--compression", "none"]
    plain, reloaded = tmp_path / "plain.exr", tmp_path / "reloaded.exr"
    run_host(common + ["--render", "3", "--out", str(plain)])
    log = run_host(common + ["--render", "5", "--reload-at", "2", "--out", str(reloaded)])
    assert "Plugin reloaded before frame 2 (1 reloads, 1 ok, 0 failed" in log
    assert log.count("Loaded plugin:") == 2 and "Saving EXR with AOVs (3 samples)" in log
    assert plain.read_bytes() == reloaded.read_bytes()


def test_host_watch_reloads_a_rebuilt_library(tmp_path):
    """--watch: while the host renders, the library file it was started from
    changes (its modification time moves, as after a rebuild); the host
    reloads it between two frames and restarts the accumulation, so its image
    equals a fresh run of the frames rendered after the reload.  The host
    holds frame 3 until a trigger file appears (--wait-at / --wait-file), and
    the test moves the library's time before creating it, so the reload
    happens before frame 3 however fast either side runs."""
    import os
    import shutil
    lib = tmp_path / "libquasi_hip.so"
    shutil.copyfile(ROOT / "quasi_amd" / "libquasi_hip.so", lib)
    common = ["--width", "96", "--height", "64", "--spp-per-render", "2", "--depth", "4", "--integrator", "nee",
              "--rng", "hashed", "--seed", "9", "--compression", "none"]
    frames, at = 8, 3
    trigger = tmp_path / "go"
    watched = tmp_path / "watched.exr"
    args = [str(HOST), str(lib), "--watch", "--render", str(frames), "--out", str(watched),
            "--wait-at", str(at), "--wait-file", str(trigger)] + common
    print("[host]", " ".join(args), flush=True)
    proc = subprocess.Popen(args, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    try:
        head = []
        for line in proc.stdout:             # the host flushes this line when it holds frame `at`
            head.append(line)
            if f"Waiting before frame {at}" in line:
                break
        st = os.stat(lib)
        os.utime(lib, ns=(st.st_atime_ns, st.st_mtime_ns + 10_000_000_000))
        trigger.write_text("go")
        rest, err = proc.communicate(timeout=120)
        out = "".join(head) + rest
    finally:
        if proc.poll() is None:
            proc.kill()
            proc.wait()
    print(out[-1500:], err[-1500:], flush=True)
    assert proc.returncode == 0
    assert f"Plugin reloaded before frame {at} (1 reloads, 1 ok, 0 failed" in out
    after = frames - at
    assert f"Saving EXR with AOVs ({after} samples)" in out
    fresh = tmp_path / "fresh.exr"
    run_host([str(lib), "--render", str(after), "--out", str(fresh)] + common)
    assert watched.read_bytes() == fresh.read_bytes()

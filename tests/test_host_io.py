"""The host side: EXR writer, binary16 conversion, scene files, the C++ plugin
loader and the headless host's failure path (CPU only).

Pins (the reference's own tests check only that files are created and that
invalid data is rejected, exr_writer_test.cpp:10-99):
- binary16 facts: 1.0 -> 0x3C00, 0.5 -> 0x3800, 65504 -> 0x7BFF, 65519 ->
  0x7BFF, >= 65520 -> infinity, 2^-24 -> 0x0001, 2^-25 -> 0 (ties to even),
  NaN stays NaN; and agreement with numpy's float16 rounding (round to
  nearest even) on 200k random values over the whole range;
- EXR round trip through an independent reader (tests/exr_reader.py): every
  HALF channel equals the float16 rounding of the input, depth.Z is the FP32
  input bit for bit, for ZIP and uncompressed files;
- channel names and types of the multi-layer file (exr_writer.cpp:105-125,
  :157-163) and of the single-layer file (:54-56).
"""
import ctypes as C
import subprocess

import numpy as np
import pytest

from quasi_amd import scenes
from tests.conftest import ROOT
from tests.exr_reader import read_exr

IO = ROOT / "quasi_amd" / "libquasi_io.so"
HOST = ROOT / "quasi_amd" / "bin" / "quasi_host"


@pytest.fixture(scope="module")
def io():
    lib = C.CDLL(str(IO))
    lib.q_io_float_to_half.restype = C.c_uint16
    lib.q_io_float_to_half.argtypes = [C.c_float]
    F = C.POINTER(C.c_float)
    lib.q_io_write_exr_aov.restype = C.c_int
    lib.q_io_write_exr_aov.argtypes = [C.c_char_p, C.c_uint32, C.c_uint32, F, F, F, F, C.c_int]
    lib.q_io_write_exr_rgba.restype = C.c_int
    lib.q_io_write_exr_rgba.argtypes = [C.c_char_p, C.c_uint32, C.c_uint32, F, C.c_int]
    lib.q_io_scene_info.restype = C.c_int
    lib.q_io_scene_info.argtypes = [C.c_char_p, C.POINTER(C.c_uint32), F]
    return lib


def fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None


def test_binary16_facts(io):
    h = io.q_io_float_to_half
    assert h(1.0) == 0x3C00 and h(0.5) == 0x3800 and h(-2.0) == 0xC000
    assert h(65504.0) == 0x7BFF and h(65519.0) == 0x7BFF
    assert h(65520.0) == 0x7C00 and h(1e10) == 0x7C00 and h(-1e10) == 0xFC00 and h(float("inf")) == 0x7C00
    assert h(2.0 ** -24) == 0x0001 and h(2.0 ** -25) == 0 and h(1.5 * 2.0 ** -25) == 0x0001
    assert h(2.0 ** -14) == 0x0400 and h(0.0) == 0 and h(-0.0) == 0x8000
    v = h(float("nan"))
    assert (v & 0x7C00) == 0x7C00 and (v & 0x3FF) != 0


def test_binary16_matches_numpy_rounding(io):
    rng = np.random.Generator(np.random.PCG64(3))
    x = np.concatenate([rng.uniform(-70000, 70000, 50000), rng.normal(0, 1, 50000),
                        np.exp(rng.uniform(-20, 11, 50000)) * rng.choice([-1, 1], 50000),
                        rng.uniform(-2 ** -13, 2 ** -13, 50000)]).astype(np.float32)
    ours = np.array([io.q_io_float_to_half(float(v)) for v in x], dtype=np.uint16)
    with np.errstate(over="ignore"):
        ref = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(ours, ref)


def images(w, h, seed=0):
    g = np.random.Generator(np.random.PCG64(seed))
    b = g.uniform(0, 4, (h, w, 4)).astype(np.float32)
    b[..., 3] = 1.0
    a = g.uniform(0, 1, (h, w, 4)).astype(np.float32)
    n = g.uniform(0, 1, (h, w, 4)).astype(np.float32)
    d = np.repeat(g.uniform(0, 1, (h, w, 1)), 4, axis=2).astype(np.float32)
    return b, a, n, d


@pytest.mark.parametrize("compression", [3, 0])
@pytest.mark.parametrize("w,h", [(37, 29), (64, 16), (5, 1)])
def test_aov_exr_round_trip(io, tmp_path, compression, w, h):
    b, a, n, d = images(w, h, seed=w * h)
    path = tmp_path / "sub" / "aov.exr"
    assert io.q_io_write_exr_aov(str(path).encode(), w, h, fp(b), fp(a), fp(n), fp(d), compression) == 0
    hdr, ch = read_exr(path)
    assert hdr["compression"][1][0] == compression
    assert hdr["lineOrder"][1][0] == 0
    names = ["A", "B", "G", "R", "albedo.B", "albedo.G", "albedo.R", "depth.Z", "normal.X", "normal.Y", "normal.Z"]
    assert sorted(ch) == names
    for name, src, k in (("R", b, 0), ("G", b, 1), ("B", b, 2), ("A", b, 3), ("albedo.R", a, 0), ("albedo.G", a, 1),
                         ("albedo.B", a, 2), ("normal.X", n, 0), ("normal.Y", n, 1), ("normal.Z", n, 2)):
        assert ch[name].dtype == np.float16
        assert np.array_equal(ch[name].view(np.uint16), src[..., k].astype(np.float16).view(np.uint16)), name
    assert ch["depth.Z"].dtype == np.float32 and np.array_equal(ch["depth.Z"], d[..., 0])


def test_zip_blocks_compress_and_rgba_file(io, tmp_path):
    w, h = 256, 64
    b = np.zeros((h, w, 4), dtype=np.float32)
    b[..., 0] = np.linspace(0, 1, w, dtype=np.float32)
    b[..., 3] = 1.0
    pz, pn = tmp_path / "z.exr", tmp_path / "n.exr"
    assert io.q_io_write_exr_rgba(str(pz).encode(), w, h, fp(b), 3) == 0
    assert io.q_io_write_exr_rgba(str(pn).encode(), w, h, fp(b), 0) == 0
    assert pz.stat().st_size < pn.stat().st_size / 4
    for p in (pz, pn):
        _, ch = read_exr(p)
        assert sorted(ch) == ["A", "B", "G", "R"]
        assert np.array_equal(ch["R"], b[..., 0].astype(np.float16))


def test_invalid_data_is_rejected(io, tmp_path):
    b = np.zeros((4, 4, 4), dtype=np.float32)
    assert io.q_io_write_exr_aov(str(tmp_path / "x.exr").encode(), 4, 4, None, None, None, None, 3) == 1
    assert io.q_io_write_exr_rgba(str(tmp_path / "y.exr").encode(), 0, 4, fp(b), 3) == 1


def test_scene_file_round_trip(io, tmp_path):
    m = scenes.cornell36(0)
    p = tmp_path / "c.qscene"
    scenes.save_qscene(m, p)
    counts = (C.c_uint32 * 2)()
    cam = (C.c_float * 10)()
    assert io.q_io_scene_info(str(p).encode(), counts, cam) == 0
    assert (counts[0], counts[1]) == (m.vertices.shape[0], m.num_triangles)
    assert np.allclose(list(cam), list(m.camera["position"]) + list(m.camera["target"]) + list(m.camera["up"])
                       + [m.camera["fov"]])
    bad = tmp_path / "bad.qscene"
    bad.write_bytes(p.read_bytes()[:100])
    assert io.q_io_scene_info(str(bad).encode(), counts, cam) == 2      # bad_format


FAKE = r"""
#include <stdint.h>
typedef struct { uint32_t major, minor, patch; } V;
typedef struct { const char* name; V version; const char* description; const char* author; } Info;
static int dummy;
uint32_t Q_plugin_abi_version(void) { return ABI; }
Info Q_plugin_get_info(void) { Info i = {"fake", {1, 2, 3}, "fake plugin", "tests"}; return i; }
void* Q_plugin_create(void* ctx) { (void)ctx; return CREATE_OK ? (void*)&dummy : 0; }
void Q_plugin_destroy(void* h) { (void)h; }
void Q_plugin_update(void* h, float dt) { (void)h; (void)dt; }
#if HAVE_RENDER
void Q_plugin_render(void* h, void* f) { (void)h; (void)f; }
#endif
"""


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    d = tmp_path_factory.mktemp("probe")
    exe = d / "loader_probe"
    subprocess.run(["g++", "-std=c++20", "-O1", "-I", str(ROOT / "include"), "-I", str(ROOT / "quasi_amd" / "host"),
                    str(ROOT / "tests" / "native" / "loader_probe.cpp"), "-o", str(exe), "-ldl"], check=True)
    return exe, d


def run_probe(probe, abi=3, create_ok=1, have_render=1):
    exe, d = probe
    src = d / f"f{abi}{create_ok}{have_render}.c"
    src.write_text(FAKE)
    so = d / f"libf{abi}{create_ok}{have_render}.so"
    subprocess.run(["gcc", "-shared", "-fPIC", f"-DABI={abi}u", f"-DCREATE_OK={create_ok}",
                    f"-DHAVE_RENDER={have_render}", str(src), "-o", str(so)], check=True)
    return subprocess.run([str(exe), str(so)], capture_output=True, text=True, check=True).stdout.strip()


def test_cpp_loader_taxonomy(probe):
    """quasi_amd/host/loader.hpp behaves as the reference loader.hpp:65-234."""
    assert run_probe(probe) == "ok 3 fake readback=0 aov=0"
    assert run_probe(probe, abi=1).startswith("ok 1")
    assert run_probe(probe, abi=0) == "error loader ABI version mismatch"
    assert run_probe(probe, abi=4) == "error loader ABI version mismatch"
    assert run_probe(probe, have_render=0) == "error loader Required symbol not found"
    assert run_probe(probe, create_ok=0) == "error loader Plugin creation failed"
    exe, d = probe
    out = subprocess.run([str(exe), str(d / "missing.so")], capture_output=True, text=True).stdout.strip()
    assert out == "error dynamic_library file not found"
    bad = d / "libbad.so"
    bad.write_bytes(b"\x7fELF broken")
    out = subprocess.run([str(exe), str(bad)], capture_output=True, text=True).stdout.strip()
    assert out == "error dynamic_library failed to load library"


def test_cpp_loader_on_the_real_plugin_without_gpu(probe):
    """The real plugin resolves through the C++ loader; without a GPU its
    create fails cleanly."""
    from tests.conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is present")
    exe, _ = probe
    out = subprocess.run([str(exe), str(ROOT / "quasi_amd" / "libquasi_hip.so")], capture_output=True,
                         text=True).stdout.strip()
    assert out == "error loader Plugin creation failed"


def test_host_usage_and_missing_gpu(tmp_path):
    from tests.conftest import gpu_available
    r = subprocess.run([str(HOST), "--bogus"], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr
    if gpu_available():
        pytest.skip("a GPU is present")
    r = subprocess.run([str(HOST), "--render", "1", "--out", str(tmp_path / "x.exr")], capture_output=True, text=True)
    assert r.returncode == 1 and "HIP device not found" in r.stderr

"""The drop-in boundary: libquasi_hip.so's C ABI and the loader's error taxonomy.

Mirrors the reference's test/plugin_test.cpp (version ordering :17-38, error
strings :44-49, dynamic_library :55-86, extension helpers :92-126, symbol-name
constants :132-142) and adds what the reference never tests: real and fake
plugins through the loader (missing symbol, ABI 0 and 4, create returning
NULL, corrupt library), and the plugin's create-time failure paths.
No GPU work is launched here.
"""
import ctypes as C
import re
import subprocess

import pytest

from quasi_amd import abi
from quasi_amd.plugin import (DynamicLibrary, LibraryError, LoaderError, Plugin, PluginError,
                              ensure_library_extension, shared_library_extension)
from tests.conftest import ROOT, gpu_available

PLUGIN = ROOT / "quasi_amd" / "libquasi_hip.so"


def declared_functions():
    names = set()
    for h in (ROOT / "include" / "quasi").glob("*.h"):
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(Q_\w+)\s*\(", h.read_text(), re.M):
            names.add(m.group(1))
    return names


def test_headers_declare_the_reference_abi():
    names = declared_functions()
    assert set(abi.REQUIRED_SYMBOLS) | set(abi.OPTIONAL_SYMBOLS) <= names
    assert set(abi.EXTENSION_SYMBOLS) <= names


def test_plugin_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", str(PLUGIN)], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert declared_functions() <= exported
    # nothing else leaks out of the library (hidden visibility by default)
    assert {s for s in exported if not s.startswith("Q_")} <= {"_init", "_fini"}


def test_symbol_name_constants():
    # plugin_test.cpp:132-142
    assert abi.K_SYMBOL_ABI_VERSION == "Q_plugin_abi_version"
    assert abi.K_SYMBOL_CREATE == "Q_plugin_create"
    assert abi.K_SYMBOL_RENDER == "Q_plugin_render"
    assert abi.Q_PLUGIN_ABI_VERSION == 3


def test_header_constants_match_the_python_mirror():
    # every numeric #define of the headers that quasi_amd.abi also names has the same value
    checked = 0
    for h in (ROOT / "include" / "quasi").glob("*.h"):
        for m in re.finditer(r"^#define\s+(Q_\w+)\s+(-?(?:0x[0-9A-Fa-f]+|\d+))u?\b", h.read_text(), re.M):
            name, value = m.group(1), int(m.group(2), 0)
            if hasattr(abi, name):
                assert getattr(abi, name) == value, name
                checked += 1
    assert checked >= 20 and abi.Q_QUEUE_FLAT == 0xFFFFFFFF


LAYOUT_STRUCTS = ["Q_gpu_context", "Q_camera", "Q_render_frame", "Q_plugin_version", "Q_plugin_info",
                  "Q_plugin_context", "Q_readback_result", "Q_aov_buffer", "Q_readback_aov_result",
                  "Q_scene_desc", "Q_render_params", "Q_kernel_stat", "Q_render_stats", "Q_batch_scene",
                  "Q_batch_desc", "Q_autotune_entry", "Q_autotune_result", "Q_adjoint_desc", "Q_adjoint_result"]


def test_struct_layouts_match_c(tmp_path):
    """Every ctypes mirror has the size and field offsets the C compiler gives
    the header's struct (compiled here with gcc as C)."""
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "quasi/quasi_hip_ext.h"', "int main(void) {"]
    for s in LAYOUT_STRUCTS:
        lines.append(f'printf("{s} size %zu\\n", sizeof({s}));')
        for f, _ in getattr(abi, s)._fields_:
            lines.append(f'printf("{s} {f} %zu\\n", offsetof({s}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines():
        s, f, v = line.split()
        cls = getattr(abi, s)
        got = C.sizeof(cls) if f == "size" else getattr(cls, f).offset
        assert got == int(v), (s, f)


def test_library_errors():
    # plugin_test.cpp:44-49, 55-86
    assert LibraryError.FILE_NOT_FOUND.value == "file not found"
    assert LibraryError.LOAD_FAILED.value == "failed to load library"
    assert LibraryError.SYMBOL_NOT_FOUND.value == "symbol not found"
    assert LibraryError.NOT_LOADED.value == "library not loaded"
    with pytest.raises(PluginError) as e:
        DynamicLibrary.open("/nonexistent/library.so")
    assert e.value.error is LibraryError.FILE_NOT_FOUND
    lib = DynamicLibrary()
    assert not lib.is_loaded()
    with pytest.raises(PluginError) as e:
        lib.get_symbol("some_function")
    assert e.value.error is LibraryError.NOT_LOADED


def test_extension_helpers():
    # plugin_test.cpp:92-126
    assert shared_library_extension() == ".so"
    assert str(ensure_library_extension("libfoo")) == "libfoo.so"
    for name in ("libfoo.so", "libfoo.dylib", "libfoo.dll"):
        assert str(ensure_library_extension(name)) == name


def test_corrupt_library_fails_to_load(tmp_path):
    bad = tmp_path / "libbad.so"
    bad.write_bytes(b"\x7fELF not really a library")
    with pytest.raises(PluginError) as e:
        DynamicLibrary.open(bad)
    assert e.value.error is LibraryError.LOAD_FAILED


FAKE = r"""
#include <stdint.h>
#include <stddef.h>
typedef struct { uint32_t major, minor, patch; } V;
typedef struct { const char* name; V version; const char* description; const char* author; } Info;
static int dummy;
uint32_t Q_plugin_abi_version(void) { return ABI; }
Info Q_plugin_get_info(void) { Info i = {"fake", {1, 2, 3}, "fake plugin", "tests"}; return i; }
void* Q_plugin_create(void* ctx) { (void)ctx; return CREATE_OK ? (void*)&dummy : NULL; }
void Q_plugin_destroy(void* h) { (void)h; }
void Q_plugin_update(void* h, float dt) { (void)h; (void)dt; }
#if HAVE_RENDER
void Q_plugin_render(void* h, void* f) { (void)h; (void)f; }
#endif
"""


def build_fake(tmp_path, name, abi_version=3, create_ok=1, have_render=1):
    src = tmp_path / f"{name}.c"
    src.write_text(FAKE)
    so = tmp_path / f"lib{name}.so"
    subprocess.run(["gcc", "-shared", "-fPIC", "-O1", f"-DABI={abi_version}u", f"-DCREATE_OK={create_ok}",
                    f"-DHAVE_RENDER={have_render}", str(src), "-o", str(so)], check=True)
    return so


def fake_context():
    gpu = abi.Q_gpu_context(abi.Q_GPU_BACKEND_HIP, None, None, None)
    ctx = abi.Q_plugin_context(64, 64, None, C.pointer(gpu), abi.LOG_FN(lambda h, m: None),
                               abi.SHUTDOWN_FN(lambda h: None))
    ctx._keep = gpu
    return ctx


def test_loader_accepts_a_valid_fake(tmp_path):
    p = Plugin.load(DynamicLibrary.open(build_fake(tmp_path, "ok")), fake_context())
    assert p.is_valid() and p.abi_version() == 3
    assert p.info().name == b"fake" and p.info().version.minor == 2
    assert not p.supports_readback() and not p.supports_readback_aov()
    p.destroy()
    assert not p.is_valid()


@pytest.mark.parametrize("version", [0, 4])
def test_loader_rejects_abi_versions_outside_1_to_3(tmp_path, version):
    lib = DynamicLibrary.open(build_fake(tmp_path, f"abi{version}", abi_version=version))
    with pytest.raises(PluginError) as e:
        Plugin.load(lib, fake_context())
    assert e.value.error is LoaderError.ABI_MISMATCH


@pytest.mark.parametrize("version", [1, 2])
def test_loader_accepts_older_abi_versions(tmp_path, version):
    lib = DynamicLibrary.open(build_fake(tmp_path, f"abi{version}", abi_version=version))
    assert Plugin.load(lib, fake_context()).is_valid()


def test_loader_requires_render_symbol(tmp_path):
    lib = DynamicLibrary.open(build_fake(tmp_path, "norender", have_render=0))
    with pytest.raises(PluginError) as e:
        Plugin.load(lib, fake_context())
    assert e.value.error is LoaderError.SYMBOL_NOT_FOUND


def test_loader_reports_create_failure(tmp_path):
    lib = DynamicLibrary.open(build_fake(tmp_path, "nocreate", create_ok=0))
    with pytest.raises(PluginError) as e:
        Plugin.load(lib, fake_context())
    assert e.value.error is LoaderError.CREATE_FAILED


def hip_plugin_context(backend, messages):
    dev = C.c_int(0)
    gpu = abi.Q_gpu_context(backend, C.cast(C.pointer(dev), C.c_void_p), None, None)
    cb = abi.LOG_FN(lambda h, m: messages.append(m.decode()))
    ctx = abi.Q_plugin_context(32, 32, None, C.pointer(gpu), cb, abi.SHUTDOWN_FN(lambda h: None))
    ctx._keep = (dev, gpu, cb)
    return ctx


def test_real_plugin_metadata():
    lib = DynamicLibrary.open(PLUGIN)
    v = lib.get_symbol("Q_plugin_abi_version", C.c_uint32)
    assert v() == 3
    info = lib.get_symbol("Q_plugin_get_info", abi.Q_plugin_info)()
    assert b"HIP" in info.name and info.version.major == 1


def test_real_plugin_rejects_non_hip_context():
    msgs = []
    lib = DynamicLibrary.open(PLUGIN)
    with pytest.raises(PluginError) as e:
        Plugin.load(lib, hip_plugin_context(abi.Q_GPU_BACKEND_METAL, msgs))
    assert e.value.error is LoaderError.CREATE_FAILED
    assert any("HIP" in m for m in msgs)


def test_real_plugin_null_arguments_are_safe():
    lib = DynamicLibrary.open(PLUGIN)
    create = lib.get_symbol("Q_plugin_create", C.c_void_p, [C.c_void_p])
    assert create(None) is None
    lib.get_symbol("Q_plugin_destroy", None, [C.c_void_p])(None)
    lib.get_symbol("Q_plugin_render", None, [C.c_void_p, C.c_void_p])(None, None)
    rb = lib.get_symbol("Q_plugin_readback", abi.Q_readback_result, [C.c_void_p])(None)
    assert not rb.data
    lib.get_symbol("Q_plugin_readback_free", None, [C.c_void_p])(None)
    lib.get_symbol("Q_plugin_readback_aov_free", None, [C.c_void_p])(None)
    for name in ("Q_plugin_reset_stats", "Q_plugin_synchronize"):
        assert lib.get_symbol(name, C.c_int32, [C.c_void_p])(None) == abi.Q_EXT_ERR_ARGUMENT


@pytest.mark.skipif(gpu_available(), reason="a GPU is present")
def test_real_plugin_create_without_gpu_logs_and_returns_null():
    msgs = []
    lib = DynamicLibrary.open(PLUGIN)
    with pytest.raises(PluginError) as e:
        Plugin.load(lib, hip_plugin_context(abi.Q_GPU_BACKEND_HIP, msgs))
    assert e.value.error is LoaderError.CREATE_FAILED
    assert any("no HIP device" in m for m in msgs)

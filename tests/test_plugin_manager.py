"""The host's hot-reload plugin manager (quasi_amd/host/manager.hpp, after the
reference's src/quasi/plugin/manager.hpp:78-330 and
src/quasi/async/file_watcher.hpp:28-133) on the CPU, with a GPU-free plugin
(tests/native/fake_plugin.c) built in several variants: loading from a private
temporary copy, reload on a change of the library file (old handle destroyed
before the new one is created, hooks in order, statistics), a failed reload
(broken file, wrong ABI, missing symbol, refused create) that leaves nothing
loaded and recovers on the next good build, and removal of the copy."""
import ctypes as C
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
NATIVE = ROOT / "tests" / "native"

# manager error codes (Q::plugin::manager::error), as returned by the probe (1 + code)
FILE_NOT_FOUND, LOAD_FAILED, ABI_MISMATCH, SYMBOL_MISSING, CREATE_FAILED = 1, 2, 3, 4, 5


@pytest.fixture(scope="module")
def built(tmp_path_factory):
    d = tmp_path_factory.mktemp("manager")
    variants = {"v1": ["-DFAKE_VERSION=1"], "v2": ["-DFAKE_VERSION=2"], "abi4": ["-DFAKE_ABI=4"],
                "norender": ["-DFAKE_NO_RENDER"], "nocreate": ["-DFAKE_FAIL_CREATE"]}
    libs = {}
    for name, flags in variants.items():
        out = d / f"libfake_{name}.so"
        subprocess.run(["gcc", "-O2", "-shared", "-fPIC", "-I", str(ROOT / "include"), *flags,
                        str(NATIVE / "fake_plugin.c"), "-o", str(out)], check=True)
        libs[name] = out
    probe = d / "libmgrprobe.so"
    subprocess.run(["g++", "-std=c++20", "-O2", "-shared", "-fPIC", "-I", str(ROOT / "include"),
                    "-I", str(ROOT / "quasi_amd" / "host"), str(NATIVE / "manager_probe.cpp"), "-o", str(probe)],
                   check=True)
    lib = C.CDLL(str(probe))
    lib.mgr_new.restype = C.c_void_p
    lib.mgr_new.argtypes = [C.c_char_p, C.c_uint, C.c_uint, C.c_uint]
    for f in ("mgr_free",):
        getattr(lib, f).argtypes = [C.c_void_p]
    for f in ("mgr_load", "mgr_reload", "mgr_poll", "mgr_is_loaded", "mgr_version"):
        getattr(lib, f).argtypes = [C.c_void_p]
        getattr(lib, f).restype = C.c_int
    lib.mgr_render.argtypes = [C.c_void_p, C.c_int]
    lib.mgr_readback.argtypes = [C.c_void_p]
    lib.mgr_readback.restype = C.c_float
    lib.mgr_stats.argtypes = [C.c_void_p, C.POINTER(C.c_ulonglong), C.POINTER(C.c_float)]
    lib.mgr_copy_path.argtypes = [C.c_void_p, C.c_char_p, C.c_uint]
    lib.mgr_events.argtypes = [C.c_void_p, C.c_char_p, C.c_uint]
    return lib, libs, d


class Managed:
    def __init__(self, lib, path):
        self.lib = lib
        self.h = lib.mgr_new(str(path).encode(), 8, 4, 0)

    def events(self):
        b = C.create_string_buffer(4096)
        self.lib.mgr_events(self.h, b, 4096)
        return [e for e in b.value.decode().split("\n") if e]

    def stats(self):
        c = (C.c_ulonglong * 3)()
        t = C.c_float()
        self.lib.mgr_stats(self.h, c, C.byref(t))
        return list(c), t.value

    def copy_path(self):
        b = C.create_string_buffer(1024)
        self.lib.mgr_copy_path(self.h, b, 1024)
        return Path(b.value.decode()) if b.value else None

    def close(self):
        self.lib.mgr_free(self.h)


def install(src, dst, bump=0):
    """Copies a build over the watched file and moves its modification time
    forward by `bump` seconds, so the change is seen whatever the clock's
    resolution."""
    shutil.copyfile(src, dst)
    st = os.stat(dst)
    os.utime(dst, ns=(st.st_atime_ns, st.st_mtime_ns + bump * 1_000_000_000))


def test_load_renders_from_a_private_copy(built, tmp_path):
    lib, libs, _ = built
    target = tmp_path / "libplugin.so"
    install(libs["v1"], target)
    m = Managed(lib, target)
    assert lib.mgr_load(m.h) == 0 and lib.mgr_is_loaded(m.h) == 1
    copy = m.copy_path()
    assert copy is not None and copy.exists() and copy != target
    assert copy.name.startswith("quasi_plugin_") and copy.name.endswith("_libplugin.so")
    assert lib.mgr_version(m.h) == 1
    lib.mgr_render(m.h, 3)
    assert lib.mgr_readback(m.h) == 1003.0
    assert m.events() == ["fake v1 create 0"]
    assert lib.mgr_poll(m.h) == 0                # nothing changed
    m.close()
    assert not copy.exists()                     # the copy goes with the manager


def test_reload_on_change(built, tmp_path):
    lib, libs, _ = built
    target = tmp_path / "libplugin.so"
    install(libs["v1"], target)
    m = Managed(lib, target)
    assert lib.mgr_load(m.h) == 0
    first = m.copy_path()
    lib.mgr_render(m.h, 2)
    m.events()
    install(libs["v2"], target, bump=5)          # a rebuild lands on the watched path
    assert lib.mgr_poll(m.h) == 1
    assert m.events() == ["pre_unload", "fake v1 destroy 2", "fake v2 create 0", "post_load"]
    assert lib.mgr_version(m.h) == 2 and lib.mgr_is_loaded(m.h) == 1
    assert not first.exists() and m.copy_path().exists()
    lib.mgr_render(m.h, 1)
    assert lib.mgr_readback(m.h) == 2001.0       # the new build, from a fresh state
    counts, secs = m.stats()
    assert counts == [1, 1, 0] and secs >= 0.0
    assert lib.mgr_poll(m.h) == 0
    m.close()


@pytest.mark.parametrize("bad,code", [("garbage", LOAD_FAILED), ("abi4", ABI_MISMATCH),
                                      ("norender", SYMBOL_MISSING), ("nocreate", CREATE_FAILED)])
def test_failed_reload_leaves_nothing_loaded_then_recovers(built, tmp_path, bad, code):
    lib, libs, d = built
    target = tmp_path / "libplugin.so"
    install(libs["v1"], target)
    m = Managed(lib, target)
    assert lib.mgr_load(m.h) == 0
    m.events()
    if bad == "garbage":
        target.write_bytes(b"not a shared library")
        st = os.stat(target)
        os.utime(target, ns=(st.st_atime_ns, st.st_mtime_ns + 5_000_000_000))
    else:
        install(libs[bad], target, bump=5)
    assert lib.mgr_reload(m.h) == code
    ev = m.events()
    assert ev[0] == "pre_unload" and ev[1] == "fake v1 destroy 0" and ev[-1].startswith("error: reload failed")
    assert "post_load" not in ev
    assert lib.mgr_is_loaded(m.h) == 0 and lib.mgr_version(m.h) == -1
    lib.mgr_render(m.h, 2)                       # no plugin: a no-op
    assert lib.mgr_readback(m.h) == -1.0
    assert lib.mgr_poll(m.h) == 0                # the failed file is not retried until it changes
    install(libs["v2"], target, bump=10)
    assert lib.mgr_poll(m.h) == 1
    assert lib.mgr_version(m.h) == 2 and lib.mgr_is_loaded(m.h) == 1
    counts, _ = m.stats()
    assert counts == [2, 1, 1]
    m.close()


def test_missing_library(built, tmp_path):
    lib, _, _ = built
    m = Managed(lib, tmp_path / "libnothing.so")
    assert lib.mgr_load(m.h) == FILE_NOT_FOUND and lib.mgr_is_loaded(m.h) == 0
    assert lib.mgr_poll(m.h) == 0
    m.close()


def test_manager_source_needs_no_gpu_headers():
    """The manager and watcher build from the C ABI headers alone (the probe
    above compiles with g++ and no ROCm include path)."""
    for f in ("manager.hpp", "file_watcher.hpp", "loader.hpp", "dynamic_library.hpp"):
        text = (ROOT / "quasi_amd" / "host" / f).read_text()
        assert "hip_runtime" not in text and "hip/" not in text, f

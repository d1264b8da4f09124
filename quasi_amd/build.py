"""Builds the native parts in-tree.

- ``quasi_amd/libquasi_hip.so``: the plugin (HIP kernels for gfx950 + the
  Q_plugin_* C ABI), compiled with hipcc.
- ``quasi_amd/bin/quasi_host``: the headless C++ host that dlopens a plugin
  exactly like the reference host (src/quasi/host/main.cpp).

The CPU oracle used by the tests is built separately by ``oracle/Makefile``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
HOST = PKG / "host"
BUILD = ROOT / "_build"
PLUGIN_SO = PKG / "libquasi_hip.so"
HOST_BIN = PKG / "bin" / "quasi_host"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = ["-O3", "-std=c++20", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter"]
# Device code: gfx950 only, packed-FP32 instruction selection disabled.
DEVICE = ["-x", "hip", f"--offload-arch={ARCH}", "-fvisibility=hidden",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HOSTONLY = ["-x", "c++", "-fvisibility=hidden", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]

PLUGIN_SOURCES = [
    ("kernels_mega.hip", DEVICE),
    ("kernels_wave.hip", DEVICE),
    ("kernels_paths.hip", DEVICE),
    ("kernels_util.hip", DEVICE),
    ("kernels_adjoint.hip", DEVICE),
    ("bvh_gpu.hip", DEVICE),
    ("bvh_layout.hip", DEVICE),
    ("plugin_noalloc.hip", DEVICE),
    ("plugin.cpp", HOSTONLY),
    ("plugin_batch.cpp", HOSTONLY),
    ("plugin_adjoint.cpp", HOSTONLY),
    ("bvh_build.cpp", HOSTONLY),
    ("scene_cornell.cpp", HOSTONLY),
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(str(c) for c in cmd), flush=True)
    r = subprocess.run([str(c) for c in cmd], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(map(str, cmd))}\n{r.stdout}\n{r.stderr}")
    return r


def build_plugin(verbose: bool = False, jobs: int = 8) -> Path:
    BUILD.mkdir(exist_ok=True)
    objs = []
    cmds = []
    for src, flags in PLUGIN_SOURCES:
        obj = BUILD / (src + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *COMMON, *flags, "-I", ROOT / "include", "-c", CSRC / src, "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    tmp = PLUGIN_SO.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-pthread"], verbose)
    os.replace(tmp, PLUGIN_SO)
    return PLUGIN_SO


HOST_BIN_SOURCES = ["main.cpp", "hip_context.cpp", "exr_writer.cpp", "scene_file.cpp"]
IO_LIB = PKG / "libquasi_io.so"
IO_LIB_SOURCES = ["exr_writer.cpp", "scene_file.cpp", "io_capi.cpp"]


def build_host(verbose: bool = False) -> Path | None:
    """The headless host (quasi_amd/bin/quasi_host: HIP context, plugin loader,
    EXR writer) and libquasi_io.so (EXR writer and scene-file reader for the
    tests)."""
    if not (HOST / "main.cpp").exists():
        return None
    HOST_BIN.parent.mkdir(exist_ok=True)
    tmp = HOST_BIN.with_suffix(".tmp")
    _run([HIPCC, "-O2", "-std=c++20", "-Wall", "-Wextra", "-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
          "-I", ROOT / "include", "-I", HOST, *[str(HOST / s) for s in HOST_BIN_SOURCES], "-o", tmp,
          "-ldl", "-lz", "-pthread"], verbose)
    os.replace(tmp, HOST_BIN)
    tmp = IO_LIB.with_suffix(".so.tmp")
    _run(["g++", "-O2", "-std=c++20", "-Wall", "-Wextra", "-fPIC", "-shared", "-fvisibility=hidden",
          "-I", ROOT / "include", "-I", HOST, *[str(HOST / s) for s in IO_LIB_SOURCES], "-o", tmp, "-lz",
          "-pthread"], verbose)
    os.replace(tmp, IO_LIB)
    return HOST_BIN


LAYOUT_PROBE = ROOT / "tests" / "native" / "liblayoutprobe.so"


def build_test_probes(verbose: bool = False) -> Path:
    """tests/native/liblayoutprobe.so: the GPU scene build's layout next to the
    host converter, from the plugin's own objects (build_plugin first)."""
    obj = BUILD / "layout_probe.cpp.o"
    _run([HIPCC, *COMMON, *HOSTONLY, "-I", ROOT / "include", "-I", CSRC, "-c",
          ROOT / "tests" / "native" / "layout_probe.cpp", "-o", obj], verbose)
    parts = [BUILD / "bvh_gpu.hip.o", BUILD / "bvh_layout.hip.o", BUILD / "bvh_build.cpp.o"]
    tmp = LAYOUT_PROBE.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, obj, *parts, "-pthread"], verbose)
    os.replace(tmp, LAYOUT_PROBE)
    return LAYOUT_PROBE


def build_oracle(verbose: bool = False) -> None:
    """The checker used by tests (oracle/liboracle.so and, when the reference
    sources are present, oracle/_ref)."""
    _run(["make", "-s", "-C", ROOT / "oracle", "-B", "liboracle.so"], verbose)
    if Path("/root/reference/src/quasi/scene/cornell_box.hpp").exists():
        _run(["make", "-s", "-C", ROOT / "oracle", "-B", "ref"], verbose)


def build_all(verbose: bool = False) -> None:
    if shutil.which(HIPCC) is None and not Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    build_plugin(verbose)
    build_test_probes(verbose)
    build_host(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print("built", PLUGIN_SO)

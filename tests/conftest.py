"""Shared pytest configuration.

Markers: ``gpu`` -- needs an AMD GPU (run with ``-m gpu``).  Everything else
runs on the CPU.  The native libraries are loaded as built; they are built
here only when missing.
"""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (gfx950)")


def _ensure_built():
    if not (ROOT / "oracle" / "liboracle.so").exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "liboracle.so"], check=True)
    if not (ROOT / "quasi_amd" / "libquasi_hip.so").exists():
        from quasi_amd import build
        build.build_plugin()
    if not (ROOT / "tests" / "native" / "liblayoutprobe.so").exists():
        from quasi_amd import build
        if not (build.BUILD / "bvh_layout.hip.o").exists():
            build.build_plugin()
        build.build_test_probes()


_ensure_built()


@pytest.fixture(scope="session")
def root() -> Path:
    return ROOT


@pytest.fixture(scope="session")
def golden() -> Path:
    return GOLDEN


def gpu_available() -> bool:
    """True when the AMD kernel driver's compute device node exists (checked
    without initialising any GPU runtime)."""
    return Path("/dev/kfd").exists()


def pytest_collection_modifyitems(config, items):
    # GPU tests fail loudly when selected without a GPU (no silent skip), so a
    # `-m gpu` run on a machine without a device is visibly red.
    pass

"""quasi_amd: an MI355X-native (gfx950, HIP) rendering backend for the Quasi
plugin host.

The product is ``libquasi_hip.so`` (HIP kernels + the Q_plugin_* C ABI, built
by ``quasi_amd.build``).  The Python modules here are thin host-side helpers:
``abi`` (ctypes mirrors of include/quasi/*.h), ``plugin`` (a loader mirroring
the reference's C++ loader, plus a convenience renderer) and ``scenes``
(synthetic scene generators).
"""
from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent
PLUGIN_PATH = PACKAGE_DIR / "libquasi_hip.so"

__all__ = ["PACKAGE_DIR", "PLUGIN_PATH"]

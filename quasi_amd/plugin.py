"""Host-side loader for Quasi plugins, in Python.

Mirrors the reference's C++ plugin runtime so the tests read like its own:
- ``DynamicLibrary`` follows src/quasi/plugin/dynamic_library.hpp:100-147
  (dlopen with RTLD_NOW | RTLD_LOCAL, typed dlsym, error enum at :19-24);
- ``Plugin.load`` follows src/quasi/plugin/loader.hpp:65-137 (six required
  symbols, ABI version in [1, k_plugin_abi_version], four optional readback
  symbols, then Q_plugin_create; error enum at :38-42).

``HipRenderer`` is a convenience layer over the same ABI plus the extension
exports of include/quasi/quasi_hip_ext.h (scene upload, parameters, primary
ids, statistics), returning numpy arrays.
"""
from __future__ import annotations

import ctypes as C
import enum
import os
from pathlib import Path
from typing import Callable, Optional, Sequence

import numpy as np

from . import abi

PLUGIN_PATH = Path(__file__).resolve().parent / "libquasi_hip.so"
RTLD_NOW = 0x2
RTLD_LOCAL = 0x0


class LibraryError(enum.Enum):
    """dynamic_library.hpp:19-24."""
    FILE_NOT_FOUND = "file not found"
    LOAD_FAILED = "failed to load library"
    SYMBOL_NOT_FOUND = "symbol not found"
    NOT_LOADED = "library not loaded"


class LoaderError(enum.Enum):
    """loader.hpp:38-42 (strings from loader.hpp:288-295)."""
    ABI_MISMATCH = "ABI version mismatch"
    SYMBOL_NOT_FOUND = "Required symbol not found"
    CREATE_FAILED = "Plugin creation failed"


class PluginError(RuntimeError):
    def __init__(self, error):
        super().__init__(error.value)
        self.error = error


def shared_library_extension() -> str:
    """dynamic_library.hpp:27-30 on Linux."""
    return ".so"


def ensure_library_extension(path) -> Path:
    """dynamic_library.hpp:36-48."""
    p = Path(path)
    if p.suffix in (".so", ".dylib", ".dll"):
        return p
    return Path(str(p) + shared_library_extension())


class DynamicLibrary:
    """RAII-style dlopen wrapper (dynamic_library.hpp:56-172)."""

    def __init__(self):
        self._lib: Optional[C.CDLL] = None
        self.path: Optional[Path] = None

    @classmethod
    def open(cls, path) -> "DynamicLibrary":
        p = Path(path)
        if not p.exists():
            raise PluginError(LibraryError.FILE_NOT_FOUND)
        try:
            lib = C.CDLL(str(p), mode=RTLD_NOW | RTLD_LOCAL)
        except OSError:
            raise PluginError(LibraryError.LOAD_FAILED) from None
        d = cls()
        d._lib = lib
        d.path = p
        return d

    def is_loaded(self) -> bool:
        return self._lib is not None

    def get_symbol(self, name: str, restype=None, argtypes: Sequence = ()):
        if self._lib is None:
            raise PluginError(LibraryError.NOT_LOADED)
        try:
            fn = getattr(self._lib, name)
        except AttributeError:
            raise PluginError(LibraryError.SYMBOL_NOT_FOUND) from None
        fn.restype = restype
        fn.argtypes = list(argtypes)
        return fn

    def has_symbol(self, name: str) -> bool:
        try:
            self.get_symbol(name)
            return True
        except PluginError:
            return False

    def close(self) -> None:
        if self._lib is not None:
            import _ctypes
            _ctypes.dlclose(self._lib._handle)
            self._lib = None
            self.path = None


_H = C.c_void_p  # Q_plugin_handle*


class Plugin:
    """Typed wrapper over a loaded plugin (loader.hpp:35-283)."""

    def __init__(self):
        self.handle: Optional[int] = None
        self.fn = {}
        self.context = None

    @classmethod
    def load(cls, library: DynamicLibrary, context: abi.Q_plugin_context) -> "Plugin":
        p = cls()
        sig = {
            abi.K_SYMBOL_ABI_VERSION: (C.c_uint32, []),
            abi.K_SYMBOL_GET_INFO: (abi.Q_plugin_info, []),
            abi.K_SYMBOL_CREATE: (_H, [C.POINTER(abi.Q_plugin_context)]),
            abi.K_SYMBOL_DESTROY: (None, [_H]),
            abi.K_SYMBOL_UPDATE: (None, [_H, C.c_float]),
            abi.K_SYMBOL_RENDER: (None, [_H, C.POINTER(abi.Q_render_frame)]),
        }
        for name in abi.REQUIRED_SYMBOLS:
            try:
                p.fn[name] = library.get_symbol(name, *sig[name])
            except PluginError:
                raise PluginError(LoaderError.SYMBOL_NOT_FOUND) from None
        version = p.fn[abi.K_SYMBOL_ABI_VERSION]()
        if version < 1 or version > abi.Q_PLUGIN_ABI_VERSION:
            raise PluginError(LoaderError.ABI_MISMATCH)
        opt = {
            abi.K_SYMBOL_READBACK: (abi.Q_readback_result, [_H]),
            abi.K_SYMBOL_READBACK_FREE: (None, [C.POINTER(abi.Q_readback_result)]),
            abi.K_SYMBOL_READBACK_AOV: (abi.Q_readback_aov_result, [_H]),
            abi.K_SYMBOL_READBACK_AOV_FREE: (None, [C.POINTER(abi.Q_readback_aov_result)]),
        }
        for name, (rt, at) in opt.items():
            if library.has_symbol(name):
                p.fn[name] = library.get_symbol(name, rt, at)
        p.context = context
        handle = p.fn[abi.K_SYMBOL_CREATE](C.byref(context))
        if not handle:
            raise PluginError(LoaderError.CREATE_FAILED)
        p.handle = handle
        return p

    def update(self, dt: float) -> None:
        if self.handle:
            self.fn[abi.K_SYMBOL_UPDATE](self.handle, dt)

    def render(self, frame: abi.Q_render_frame) -> None:
        if self.handle:
            self.fn[abi.K_SYMBOL_RENDER](self.handle, C.byref(frame))

    def supports_readback(self) -> bool:
        return abi.K_SYMBOL_READBACK in self.fn and abi.K_SYMBOL_READBACK_FREE in self.fn

    def readback(self) -> abi.Q_readback_result:
        if self.handle and abi.K_SYMBOL_READBACK in self.fn:
            return self.fn[abi.K_SYMBOL_READBACK](self.handle)
        return abi.Q_readback_result()

    def readback_free(self, r: abi.Q_readback_result) -> None:
        if abi.K_SYMBOL_READBACK_FREE in self.fn:
            self.fn[abi.K_SYMBOL_READBACK_FREE](C.byref(r))

    def supports_readback_aov(self) -> bool:
        return abi.K_SYMBOL_READBACK_AOV in self.fn and abi.K_SYMBOL_READBACK_AOV_FREE in self.fn

    def readback_aov(self) -> abi.Q_readback_aov_result:
        if self.handle and abi.K_SYMBOL_READBACK_AOV in self.fn:
            return self.fn[abi.K_SYMBOL_READBACK_AOV](self.handle)
        return abi.Q_readback_aov_result()

    def readback_aov_free(self, r: abi.Q_readback_aov_result) -> None:
        if abi.K_SYMBOL_READBACK_AOV_FREE in self.fn:
            self.fn[abi.K_SYMBOL_READBACK_AOV_FREE](C.byref(r))

    def info(self) -> abi.Q_plugin_info:
        return self.fn[abi.K_SYMBOL_GET_INFO]()

    def abi_version(self) -> int:
        return int(self.fn[abi.K_SYMBOL_ABI_VERSION]())

    def is_valid(self) -> bool:
        return bool(self.handle)

    def destroy(self) -> None:
        if self.handle:
            self.fn[abi.K_SYMBOL_DESTROY](self.handle)
            self.handle = None


def scene_desc(mesh):
    """Q_scene_desc over a mesh's arrays, plus the arrays it points into (keep
    them alive for the call)."""
    v = np.ascontiguousarray(mesh.vertices, dtype=np.float32)
    i = np.ascontiguousarray(mesh.indices, dtype=np.uint32)
    a = np.ascontiguousarray(mesh.albedo, dtype=np.float32)
    e = np.ascontiguousarray(mesh.emission, dtype=np.float32)
    d = abi.Q_scene_desc(C.sizeof(abi.Q_scene_desc), v.size // 3, i.size // 3, 0,
                         v.ctypes.data_as(C.POINTER(C.c_float)), i.ctypes.data_as(C.POINTER(C.c_uint32)),
                         a.ctypes.data_as(C.POINTER(C.c_float)), e.ctypes.data_as(C.POINTER(C.c_float)))
    return d, (v, i, a, e)


def _array(buf, w, h, ch=4) -> np.ndarray:
    return np.ctypeslib.as_array(buf, shape=(h, w, ch)).copy()


class HipRenderer:
    """Owns a HIP plugin instance for scripts and tests.

    ``device`` is the HIP ordinal handed to the plugin through
    Q_gpu_context.device; ``queue`` (a hipStream_t as int) may be None, in
    which case the plugin creates its own stream.
    """

    def __init__(self, width: int, height: int, device: int = 0, queue: Optional[int] = None,
                 library: Optional[Path] = None, log: Optional[Callable[[str], None]] = None):
        path = Path(library) if library else PLUGIN_PATH
        if not path.exists():
            raise RuntimeError(f"HIP plugin not built: {path} (run quasi_amd.build)")
        self.messages: list[str] = []
        self._log_user = log
        self._device = C.c_int(device)
        self._gpu = abi.Q_gpu_context(abi.Q_GPU_BACKEND_HIP, C.cast(C.pointer(self._device), C.c_void_p),
                                      C.c_void_p(queue) if queue else None, None)

        def _log(_host, msg):
            s = msg.decode(errors="replace")
            self.messages.append(s)
            if self._log_user:
                self._log_user(s)

        self._log_cb = abi.LOG_FN(_log)
        self._shutdown_cb = abi.SHUTDOWN_FN(lambda _h: None)
        self.context = abi.Q_plugin_context(width, height, None, C.pointer(self._gpu), self._log_cb,
                                            self._shutdown_cb)
        self.lib = DynamicLibrary.open(path)
        self.plugin = Plugin.load(self.lib, self.context)
        self._bind_ext()
        self.width, self.height = width, height
        self.camera = default_camera()
        self._num_vertices = 64          # the plugin's default scene: 16 quads of 4 vertices
        self._camera_dirty = True

    def _bind_ext(self):
        g = self.lib.get_symbol
        P = C.POINTER
        self._set_scene = g("Q_plugin_set_scene", C.c_int32, [_H, P(abi.Q_scene_desc)])
        self._get_params = g("Q_plugin_get_params", C.c_int32, [_H, P(abi.Q_render_params)])
        self._set_params = g("Q_plugin_set_params", C.c_int32, [_H, P(abi.Q_render_params)])
        self._readback_ids = g("Q_plugin_readback_ids", C.c_int32, [_H, P(C.c_uint32), C.c_uint64])
        self._readback_display = g("Q_plugin_readback_display", C.c_int32, [_H, P(C.c_uint8), C.c_uint64])
        self._get_stats = g("Q_plugin_get_stats", C.c_int32, [_H, P(abi.Q_render_stats)])
        self._reset_stats = g("Q_plugin_reset_stats", C.c_int32, [_H])
        self._sync = g("Q_plugin_synchronize", C.c_int32, [_H])
        self._set_batch = g("Q_plugin_set_batch", C.c_int32, [_H, P(abi.Q_batch_desc)])
        self._render_batch = g("Q_plugin_render_batch", C.c_int32, [_H, C.c_uint32, C.c_void_p])
        self._readback_batch = g("Q_plugin_readback_batch", C.c_int32, [_H, C.c_uint32, P(abi.Q_readback_aov_result)])
        self._readback_batch_ids = g("Q_plugin_readback_batch_ids", C.c_int32,
                                     [_H, C.c_uint32, P(C.c_uint32), C.c_uint64])
        self._autotune = g("Q_plugin_autotune", C.c_int32, [_H, C.c_uint32, C.c_uint32, C.c_uint32,
                                                            P(abi.Q_autotune_result)])
        self._adjoint = g("Q_plugin_adjoint", C.c_int32, [_H, P(abi.Q_adjoint_desc), P(abi.Q_adjoint_result)])
        self._render_noalloc = g("Q_plugin_render_noalloc", C.c_int32,
                                 [_H, P(abi.Q_camera), C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p,
                                  C.c_void_p])

    def _check(self, rc: int, what: str) -> None:
        if rc != abi.Q_EXT_OK:
            tail = self.messages[-1] if self.messages else ""
            raise RuntimeError(f"{what} failed ({rc}): {tail}")

    # -- extension calls -------------------------------------------------
    def set_scene(self, mesh) -> None:
        d, keep = scene_desc(mesh)
        self._check(self._set_scene(self.plugin.handle, C.byref(d)), "set_scene")
        self._num_vertices = int(d.num_vertices)
        del keep

    def set_batch(self, scenes) -> None:
        """scenes: iterable of (mesh, width, height, seed)."""
        scenes = list(scenes)
        arr = (abi.Q_batch_scene * len(scenes))()
        keep = []
        for k, (mesh, w, h, seed) in enumerate(scenes):
            d, kp = scene_desc(mesh)
            keep.append(kp)
            c = mesh.camera
            arr[k] = abi.Q_batch_scene(w, h, seed, 0, make_camera(c["position"], c["target"], c["up"], c["fov"]), d)
        desc = abi.Q_batch_desc(C.sizeof(abi.Q_batch_desc), len(scenes), arr)
        self._check(self._set_batch(self.plugin.handle, C.byref(desc)), "set_batch")
        self.batch_sizes = [(w, h) for (_, w, h, _) in scenes]

    def render_batch(self, spp: int, stream: Optional[int] = None) -> None:
        self._check(self._render_batch(self.plugin.handle, spp, C.c_void_p(stream) if stream else None),
                    "render_batch")

    def readback_batch(self, scene: int) -> dict:
        r = abi.Q_readback_aov_result()
        rc = self._readback_batch(self.plugin.handle, scene, C.byref(r))
        out = {}
        try:
            self._check(rc, "readback_batch")
            for k, name in enumerate(("beauty", "albedo", "normal", "depth")):
                b = r.buffers[k]
                out[name] = _array(b.data, b.width, b.height, b.channels) if b.data else None
        finally:
            self.plugin.readback_aov_free(r)
        return out

    def batch_ids(self, scene: int) -> np.ndarray:
        w, h = self.batch_sizes[scene]
        out = np.empty(w * h, dtype=np.uint32)
        self._check(self._readback_batch_ids(self.plugin.handle, scene, out.ctypes.data_as(C.POINTER(C.c_uint32)),
                                             out.size), "readback_batch_ids")
        return out.reshape(h, w)

    def autotune(self, target: int = abi.Q_AUTOTUNE_BATCH, probe_spp: int = 1, flags: int = 0) -> abi.Q_autotune_result:
        r = abi.Q_autotune_result()
        r.struct_size = C.sizeof(abi.Q_autotune_result)
        self._check(self._autotune(self.plugin.handle, target, probe_spp, flags, C.byref(r)), "autotune")
        return r

    def adjoint(self, image, spp: int, sample_begin: int = 0, mode: int = abi.Q_ADJOINT_IMAGE_GRADIENT,
                seed: Optional[int] = None, num_vertices: Optional[int] = None, max_pass_paths: int = 0,
                albedo_param=None, num_albedo_params: int = 0):
        """Q_plugin_adjoint for the current scene, camera and resolution.
        image: (H, W, 4) float32 dL/dI (or the L2 target).  Returns
        (grad_albedo (T, 3) float64 -- (num_albedo_params, 3) with an albedo
        parameter map --, grad_vertices (V, 3) float64, result)."""
        img = np.ascontiguousarray(image, dtype=np.float32)
        nt = self.stats().num_triangles
        nv = self._num_vertices if num_vertices is None else num_vertices
        pmap = None
        if albedo_param is not None:
            pmap = np.ascontiguousarray(albedo_param, dtype=np.uint32)
            ga = np.zeros((num_albedo_params, 3), dtype=np.float64)
        else:
            ga = np.zeros((nt, 3), dtype=np.float64)
        gv = np.zeros((nv, 3), dtype=np.float64)
        d = abi.Q_adjoint_desc(C.sizeof(abi.Q_adjoint_desc), mode, spp, sample_begin, 0 if seed is None else seed,
                               0 if seed is None else abi.Q_ADJOINT_SEED,
                               img.ctypes.data_as(C.POINTER(C.c_float)),
                               ga.ctypes.data_as(C.POINTER(C.c_double)), gv.ctypes.data_as(C.POINTER(C.c_double)),
                               max_pass_paths, num_albedo_params if pmap is not None else 0,
                               pmap.ctypes.data_as(C.POINTER(C.c_uint32)) if pmap is not None else None)
        res = abi.Q_adjoint_result()
        res.struct_size = C.sizeof(abi.Q_adjoint_result)
        self._check(self._adjoint(self.plugin.handle, C.byref(d), C.byref(res)), "adjoint")
        return ga, gv, res

    def render_noalloc(self, sums_ptr: int, width: int, height: int, spp: int, sample_begin: int = 0,
                       camera: Optional[abi.Q_camera] = None, stream: Optional[int] = None) -> int:
        """Q_plugin_render_noalloc into the device buffer at sums_ptr (4 planes of
        width*height RGBA32F sums); returns the Q_EXT_* code."""
        cam = self.camera if camera is None else camera
        return int(self._render_noalloc(self.plugin.handle, C.byref(cam), width, height, sample_begin, spp,
                                        C.c_void_p(sums_ptr), C.c_void_p(stream) if stream else None))

    def get_params(self) -> abi.Q_render_params:
        p = abi.Q_render_params()
        p.struct_size = C.sizeof(abi.Q_render_params)
        self._check(self._get_params(self.plugin.handle, C.byref(p)), "get_params")
        return p

    def set_params(self, **kw) -> abi.Q_render_params:
        p = self.get_params()
        for k, v in kw.items():
            setattr(p, k, int(v))
        self._check(self._set_params(self.plugin.handle, C.byref(p)), "set_params")
        return p

    def stats(self) -> abi.Q_render_stats:
        s = abi.Q_render_stats()
        s.struct_size = C.sizeof(abi.Q_render_stats)
        self._check(self._get_stats(self.plugin.handle, C.byref(s)), "get_stats")
        return s

    def reset_stats(self) -> None:
        self._check(self._reset_stats(self.plugin.handle), "reset_stats")

    def synchronize(self) -> None:
        self._check(self._sync(self.plugin.handle), "synchronize")

    def primary_ids(self) -> np.ndarray:
        out = np.empty(self.width * self.height, dtype=np.uint32)
        self._check(self._readback_ids(self.plugin.handle, out.ctypes.data_as(C.POINTER(C.c_uint32)), out.size),
                    "readback_ids")
        return out.reshape(self.height, self.width)

    def readback_display(self) -> np.ndarray:
        """The 8-bit display image (Reinhard + gamma 1/2.2, plugin.mm:334-350) as
        (height, width, 4) uint8, top row first."""
        out = np.empty((self.height, self.width, 4), dtype=np.uint8)
        self._check(self._readback_display(self.plugin.handle, out.ctypes.data_as(C.POINTER(C.c_uint8)), out.size),
                    "readback_display")
        return out

    # -- reference ABI ---------------------------------------------------
    def frame(self, camera_dirty: Optional[bool] = None, stream: Optional[int] = None) -> abi.Q_render_frame:
        dirty = self._camera_dirty if camera_dirty is None else camera_dirty
        self._camera_dirty = False
        return abi.Q_render_frame(None, C.c_void_p(stream) if stream else None, self.width, self.height,
                                  self.camera, 1 if dirty else 0)

    def render(self, frames: int = 1, stream: Optional[int] = None) -> None:
        for _ in range(frames):
            f = self.frame(stream=stream)
            self.plugin.update(0.0)
            self.plugin.render(f)

    def resize(self, width: int, height: int) -> None:
        self.width, self.height = width, height

    def set_camera(self, position, target, up=(0.0, 1.0, 0.0), fov=40.0) -> None:
        self.camera = make_camera(position, target, up, fov)
        self._camera_dirty = True

    def readback(self) -> Optional[np.ndarray]:
        r = self.plugin.readback()
        if not r.data:
            return None
        try:
            return _array(r.data, r.width, r.height, r.channels)
        finally:
            self.plugin.readback_free(r)

    def readback_aov(self) -> dict:
        r = self.plugin.readback_aov()
        out = {}
        try:
            for k, name in enumerate(("beauty", "albedo", "normal", "depth")):
                b = r.buffers[k]
                out[name] = _array(b.data, b.width, b.height, b.channels) if b.data else None
        finally:
            self.plugin.readback_aov_free(r)
        return out

    def close(self) -> None:
        if getattr(self, "plugin", None) is not None:
            self.plugin.destroy()
            self.plugin = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def make_camera(position, target, up=(0.0, 1.0, 0.0), fov=40.0) -> abi.Q_camera:
    c = abi.Q_camera()
    c.position[:] = [float(x) for x in position]
    c.target[:] = [float(x) for x in target]
    c.up[:] = [float(x) for x in up]
    c.fov = float(fov)
    return c


def default_camera() -> abi.Q_camera:
    """The reference host's orbit camera at rest (main.cpp:38-45, 104-115)."""
    return make_camera((0.0, 1.0, 3.5), (0.0, 1.0, 0.0), (0.0, 1.0, 0.0), 40.0)

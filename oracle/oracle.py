"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
REF_SCENE_JSON = HERE / "_ref" / "ref_scene.json"

_lib = None


class QoParams(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in (
        "width", "height", "spp", "sample_offset", "max_depth", "rr_start", "seed", "rng_mode", "integrator",
        "use_bvh", "threads", "pixel_stride", "id_sample", "precision")] + [
        ("cam_position", C.c_float * 3), ("cam_target", C.c_float * 3), ("cam_up", C.c_float * 3),
        ("cam_fov", C.c_float)]


class QoCounts(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "camera_rays", "camera_aabb", "camera_tri", "extension_rays", "extension_aabb", "extension_tri",
        "shadow_rays", "shadow_aabb", "shadow_tri")] + [("seconds", C.c_double), ("threads", C.c_uint32),
                                                         ("pad", C.c_uint32)]


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"oracle not built: {LIB_PATH} (make -C oracle)")
        L = C.CDLL(str(LIB_PATH))
        D, U, F, P = C.c_double, C.c_uint32, C.c_float, C.POINTER
        L.qo_pcg_hash.restype = U
        L.qo_pcg_hash.argtypes = [U]
        L.qo_random_floats.argtypes = [U, U, P(F)]
        L.qo_seed.restype = U
        L.qo_seed.argtypes = [U] * 7
        L.qo_camera_direction.argtypes = [P(F), P(F), P(F), F, U, U, D, D, P(D)]
        L.qo_primary_ray.argtypes = [P(QoParams), U, U, U, P(D), P(D)]
        L.qo_cosine_sample.argtypes = [P(D), U, P(D)]
        L.qo_intersect_tri.restype = C.c_int
        L.qo_intersect_tri.argtypes = [P(D)] * 5 + [D, D, P(D)]
        L.qo_scene_create.restype = C.c_void_p
        L.qo_scene_create.argtypes = [U, U, P(F), P(U), P(F), P(F), C.c_int]
        L.qo_scene_destroy.argtypes = [C.c_void_p]
        L.qo_scene_info.argtypes = [C.c_void_p, P(D)]
        L.qo_scene_bvh.argtypes = [C.c_void_p, P(U), P(U)]
        L.qo_scene_lights.restype = U
        L.qo_scene_lights.argtypes = [C.c_void_p, P(U), P(F), P(F), P(F)]
        L.qo_closest_hits.argtypes = [C.c_void_p, U, P(D), P(D), C.c_int, P(U), P(D)]
        L.qo_render.restype = C.c_int
        L.qo_render.argtypes = [C.c_void_p, P(QoParams), P(D), P(D), P(D), P(D), P(U), P(D), P(QoCounts)]
        L.qo_hardware_concurrency.restype = U
        L.qo_adjoint.restype = C.c_int
        L.qo_adjoint.argtypes = [C.c_void_p, P(QoParams), P(F), P(D), P(D), P(D)]
        L.qo_tape_create.restype = C.c_void_p
        L.qo_tape_create.argtypes = [C.c_void_p, P(QoParams), P(F)]
        L.qo_tape_destroy.argtypes = [C.c_void_p]
        L.qo_tape_paths.restype = C.c_uint64
        L.qo_tape_paths.argtypes = [C.c_void_p]
        L.qo_tape_vertices.restype = C.c_uint64
        L.qo_tape_vertices.argtypes = [C.c_void_p]
        L.qo_tape_eval.restype = D
        L.qo_tape_eval.argtypes = [C.c_void_p, P(D), P(D), P(U), U, U]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def pcg_hash(x: int) -> int:
    return int(lib().qo_pcg_hash(x & 0xFFFFFFFF))


def random_floats(state: int, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.float32)
    lib().qo_random_floats(state & 0xFFFFFFFF, n, _p(out, C.c_float))
    return out


def seed(mode: int, x: int, y: int, w: int, h: int, sample: int, seed_value: int = 0) -> int:
    return int(lib().qo_seed(mode, x, y, w, h, sample, seed_value))


def camera_direction(camera: dict, w: int, h: int, u: float, v: float) -> np.ndarray:
    f3 = lambda t: (C.c_float * 3)(*t)
    out = np.empty(3, dtype=np.float64)
    lib().qo_camera_direction(f3(camera["position"]), f3(camera["target"]), f3(camera["up"]),
                              float(camera["fov"]), w, h, u, v, _p(out, C.c_double))
    return out


def cosine_sample(n, state: int) -> np.ndarray:
    nn = np.asarray(n, dtype=np.float64)
    out = np.empty(3, dtype=np.float64)
    lib().qo_cosine_sample(_p(nn, C.c_double), state & 0xFFFFFFFF, _p(out, C.c_double))
    return out


def intersect_tri(o, d, p0, p1, p2, tmin=1e-3, tmax=1e30):
    a = [np.asarray(x, dtype=np.float64) for x in (o, d, p0, p1, p2)]
    t = C.c_double(0)
    hit = lib().qo_intersect_tri(*[_p(x, C.c_double) for x in a], tmin, tmax, C.byref(t))
    return (t.value if hit else None)


@dataclass
class RenderResult:
    beauty: np.ndarray     # (H, W, 4) float64 sums over the samples
    albedo: np.ndarray
    normal: np.ndarray
    depth: np.ndarray
    ids: np.ndarray        # (H, W) uint32 primary ids of sample id_sample
    tprim: np.ndarray      # (H, W) float64 primary t of sample id_sample
    counts: QoCounts
    spp: int

    def mean(self, name="beauty"):
        return getattr(self, name) / float(self.spp)


class Scene:
    def __init__(self, mesh, build_bvh: bool = True):
        v = np.ascontiguousarray(mesh.vertices, dtype=np.float32)
        i = np.ascontiguousarray(mesh.indices, dtype=np.uint32)
        a = np.ascontiguousarray(mesh.albedo, dtype=np.float32)
        e = np.ascontiguousarray(mesh.emission, dtype=np.float32)
        self._keep = (v, i, a, e)
        self.handle = lib().qo_scene_create(v.size // 3, i.size // 3, _p(v, C.c_float), _p(i, C.c_uint32),
                                            _p(a, C.c_float), _p(e, C.c_float), 1 if build_bvh else 0)
        self.mesh = mesh

    def info(self) -> dict:
        out = np.zeros(6, dtype=np.float64)
        lib().qo_scene_info(self.handle, _p(out, C.c_double))
        return dict(nodes=int(out[0]), leaves=int(out[1]), depth=int(out[2]), max_leaf=int(out[3]),
                    lights=int(out[4]), build_ms=float(out[5]))

    def bvh(self):
        inf = self.info()
        nodes = np.zeros((inf["nodes"], 8), dtype=np.uint32)
        prims = np.zeros(self.mesh.num_triangles, dtype=np.uint32)
        lib().qo_scene_bvh(self.handle, _p(nodes, C.c_uint32), _p(prims, C.c_uint32))
        lo = nodes[:, 0:3].copy().view(np.float32)
        hi = nodes[:, 3:6].copy().view(np.float32)
        return lo, hi, nodes[:, 6].copy(), nodes[:, 7].copy(), prims

    def lights(self):
        n = lib().qo_scene_lights(self.handle, None, None, None, None)
        tri = np.zeros(n, dtype=np.uint32)
        cdf, psel, area = (np.zeros(n, dtype=np.float32) for _ in range(3))
        lib().qo_scene_lights(self.handle, _p(tri, C.c_uint32), _p(cdf, C.c_float), _p(psel, C.c_float),
                              _p(area, C.c_float))
        return tri, cdf, psel, area

    def closest_hits(self, origins, dirs, use_bvh=True):
        o = np.ascontiguousarray(origins, dtype=np.float64).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
        n = o.shape[0]
        tri = np.zeros(n, dtype=np.uint32)
        t = np.zeros(n, dtype=np.float64)
        lib().qo_closest_hits(self.handle, n, _p(o, C.c_double), _p(d, C.c_double), 1 if use_bvh else 0,
                              _p(tri, C.c_uint32), _p(t, C.c_double))
        return tri, t

    def render(self, width, height, spp, camera=None, sample_offset=0, max_depth=5, rr_start=3, seed=0,
               rng_mode=0, integrator=0, use_bvh=True, threads=0, pixel_stride=1, id_sample=0,
               precision=64) -> RenderResult:
        cam = camera or self.mesh.camera
        p = QoParams()
        p.width, p.height, p.spp, p.sample_offset = width, height, spp, sample_offset
        p.max_depth, p.rr_start, p.seed, p.rng_mode, p.integrator = max_depth, rr_start, seed, rng_mode, integrator
        p.use_bvh, p.threads, p.pixel_stride, p.id_sample, p.precision = (
            1 if use_bvh else 0, threads, pixel_stride, id_sample, precision)
        p.cam_position[:] = cam["position"]
        p.cam_target[:] = cam["target"]
        p.cam_up[:] = cam["up"]
        p.cam_fov = cam["fov"]
        shape = (height, width, 4)
        b, a, n, d = (np.zeros(shape, dtype=np.float64) for _ in range(4))
        ids = np.full((height, width), 0xFFFFFFFF, dtype=np.uint32)
        tp = np.zeros((height, width), dtype=np.float64)
        counts = QoCounts()
        lib().qo_render(self.handle, C.byref(p), _p(b, C.c_double), _p(a, C.c_double), _p(n, C.c_double),
                        _p(d, C.c_double), _p(ids, C.c_uint32), _p(tp, C.c_double), C.byref(counts))
        return RenderResult(b, a, n, d, ids, tp, counts, spp)

    def _params(self, width, height, spp, camera, sample_offset, max_depth, rr_start, seed, rng_mode, integrator,
                use_bvh, threads, pixel_stride, id_sample, precision) -> "QoParams":
        cam = camera or self.mesh.camera
        p = QoParams()
        p.width, p.height, p.spp, p.sample_offset = width, height, spp, sample_offset
        p.max_depth, p.rr_start, p.seed, p.rng_mode, p.integrator = max_depth, rr_start, seed, rng_mode, integrator
        p.use_bvh, p.threads, p.pixel_stride, p.id_sample, p.precision = (
            1 if use_bvh else 0, threads, pixel_stride, id_sample, precision)
        p.cam_position[:] = cam["position"]
        p.cam_target[:] = cam["target"]
        p.cam_up[:] = cam["up"]
        p.cam_fov = cam["fov"]
        return p

    def adjoint(self, width, height, spp, dimg, camera=None, sample_offset=0, max_depth=5, rr_start=3, seed=0,
                rng_mode=1, integrator=1, threads=0, pixel_stride=1, precision=64):
        """Reverse-mode gradients of l = sum_pixels dimg . (1/spp) sum_s C_s
        (dimg: (H, W, 4) float32) -> (d l/d albedo (T, 3), d l/d vertices (V, 3))."""
        p = self._params(width, height, spp, camera, sample_offset, max_depth, rr_start, seed, rng_mode, integrator,
                         True, threads, pixel_stride, 0, precision)
        dimg = np.ascontiguousarray(dimg, dtype=np.float32)
        assert dimg.shape == (height, width, 4)
        ga = np.zeros((self.mesh.num_triangles, 3), dtype=np.float64)
        gv = np.zeros((self.mesh.vertices.size // 3, 3), dtype=np.float64)
        sec = C.c_double(0)
        lib().qo_adjoint(self.handle, C.byref(p), _p(dimg, C.c_float), _p(ga, C.c_double), _p(gv, C.c_double),
                         C.byref(sec))
        self.last_adjoint_seconds = sec.value
        return ga, gv

    def tape(self, width, height, spp, dimg, camera=None, sample_offset=0, max_depth=5, rr_start=3, seed=0,
             rng_mode=1, integrator=1, pixel_stride=1) -> "Tape":
        p = self._params(width, height, spp, camera, sample_offset, max_depth, rr_start, seed, rng_mode, integrator,
                         True, 0, pixel_stride, 0, 64)
        return Tape(self, p, dimg)

    def primary_ray(self, params: dict, x, y, s):
        p = QoParams()
        cam = params.get("camera") or self.mesh.camera
        p.width, p.height = params["width"], params["height"]
        p.rng_mode, p.seed = params.get("rng_mode", 0), params.get("seed", 0)
        p.cam_position[:] = cam["position"]
        p.cam_target[:] = cam["target"]
        p.cam_up[:] = cam["up"]
        p.cam_fov = cam["fov"]
        o = np.zeros(3)
        d = np.zeros(3)
        lib().qo_primary_ray(C.byref(p), x, y, s, _p(o, C.c_double), _p(d, C.c_double))
        return o, d

    def close(self):
        if self.handle:
            lib().qo_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Tape:
    """The discrete decisions of every path of an image (FP64), for evaluating
    l = sum_pixels dimg . (1/spp) sum_s C_s at other parameters with the
    decisions frozen (central finite differences of the adjoint's estimator)."""

    def __init__(self, scene: Scene, params, dimg):
        self.scene = scene
        dimg = np.ascontiguousarray(dimg, dtype=np.float32)
        self.handle = lib().qo_tape_create(scene.handle, C.byref(params), _p(dimg, C.c_float))
        mesh = scene.mesh
        self.verts = np.ascontiguousarray(mesh.vertices, dtype=np.float64).reshape(-1, 3)
        self.albedo = np.ascontiguousarray(mesh.albedo, dtype=np.float64).reshape(-1, 3)
        self.indices = np.ascontiguousarray(mesh.indices, dtype=np.uint32).reshape(-1, 3)

    @property
    def paths(self) -> int:
        return int(lib().qo_tape_paths(self.handle))

    def eval(self, verts=None, albedo=None, tris=None, threads=0) -> float:
        v = np.ascontiguousarray(self.verts if verts is None else verts, dtype=np.float64)
        a = np.ascontiguousarray(self.albedo if albedo is None else albedo, dtype=np.float64)
        t = np.ascontiguousarray([] if tris is None else tris, dtype=np.uint32)
        return float(lib().qo_tape_eval(self.handle, _p(v, C.c_double), _p(a, C.c_double),
                                        _p(t, C.c_uint32) if t.size else None, int(t.size), threads))

    def fd(self, kind: str, index: int, comp: int, eps: float = 1e-3, richardson: bool = False) -> float:
        """Central difference of l in one parameter: kind 'albedo' (triangle
        index, channel) or 'vertex' (vertex index, axis).  richardson: the
        extrapolation (4 D(eps/2) - D(eps)) / 3, which cancels the O(eps^2)
        truncation term of the central difference D."""
        if richardson:
            return (4.0 * self.fd(kind, index, comp, eps / 2) - self.fd(kind, index, comp, eps)) / 3.0
        if kind == "albedo":
            tris = [index]
        else:
            tris = np.nonzero((self.indices == index).any(axis=1))[0]
        out = []
        for sgn in (1.0, -1.0):
            v = self.verts.copy()
            a = self.albedo.copy()
            if kind == "albedo":
                a[index, comp] += sgn * eps
            else:
                v[index, comp] += sgn * eps
            out.append(self.eval(v, a, tris))
        return (out[0] - out[1]) / (2.0 * eps)

    def close(self):
        if self.handle:
            lib().qo_tape_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def hardware_concurrency() -> int:
    return int(lib().qo_hardware_concurrency())

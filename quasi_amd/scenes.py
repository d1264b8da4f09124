"""Synthetic scene generators for the render workloads.

Every generator is deterministic for its seed (numpy PCG64 streams) and
returns a ``Mesh``: float32 vertices, uint32 triangle indices and per-triangle
float32 albedo / emission, plus the camera the workload is rendered with.

- ``cornell_ref``: the reference's own scene (src/quasi/scene/cornell_box.hpp:95-134),
  16 quads as 32 triangles.
- ``cornell36``: 36-triangle Cornell-style room (5 walls, two axis-aligned
  6-face boxes, one ceiling light of radiance 15), albedo U[0.1, 0.8] per quad.
- ``soup``: triangle soup with centroids uniform in [-1, 1]^3 and sizes
  log-uniform in [1e-3, 5e-2], albedo U[0, 1], four 0.5 x 0.5 ceiling lights.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

F32 = np.float32


@dataclass
class Mesh:
    vertices: np.ndarray          # (V, 3) float32
    indices: np.ndarray           # (T, 3) uint32
    albedo: np.ndarray            # (T, 3) float32
    emission: np.ndarray          # (T, 3) float32
    camera: dict = field(default_factory=dict)
    name: str = ""

    @property
    def num_triangles(self) -> int:
        return int(self.indices.shape[0])

    def triangles(self) -> np.ndarray:
        """(T, 3, 3) float32 vertex positions."""
        return self.vertices[self.indices]


REFERENCE_CAMERA = dict(position=(0.0, 1.0, 3.5), target=(0.0, 1.0, 0.0), up=(0.0, 1.0, 0.0), fov=40.0)


class _Builder:
    def __init__(self):
        self.v, self.i, self.a, self.e = [], [], [], []

    def quad(self, o, u, v, albedo, emission=(0.0, 0.0, 0.0)):
        """Quad (o, u, v) as triangles (o, o+u, o+u+v) and (o, o+u+v, o+v), in float32."""
        o, u, v = (np.asarray(x, dtype=F32) for x in (o, u, v))
        base = sum(len(x) for x in self.v)
        ou = o + u
        self.v.append(np.stack([o, ou, ou + v, o + v]))
        self.i.append(np.array([[0, 1, 2], [0, 2, 3]], dtype=np.uint32) + base)
        self.a.append(np.tile(np.asarray(albedo, dtype=F32), (2, 1)))
        self.e.append(np.tile(np.asarray(emission, dtype=F32), (2, 1)))

    def triangles(self, tri, albedo, emission):
        base = sum(len(x) for x in self.v)
        n = tri.shape[0]
        self.v.append(tri.reshape(-1, 3).astype(F32))
        self.i.append((np.arange(3 * n, dtype=np.uint32) + base).reshape(n, 3))
        self.a.append(albedo.astype(F32))
        self.e.append(emission.astype(F32))

    def mesh(self, camera, name) -> Mesh:
        return Mesh(np.concatenate(self.v).astype(F32), np.concatenate(self.i).astype(np.uint32),
                    np.concatenate(self.a).astype(F32), np.concatenate(self.e).astype(F32), dict(camera), name)


def _rotate_y(p, deg):
    """cornell_box.hpp:32-37, in float32."""
    rad = F32(deg) * F32(3.14159265359) / F32(180.0)
    c, s = F32(np.cos(rad)), F32(np.sin(rad))
    p = np.asarray(p, dtype=F32)
    return np.array([p[0] * c + p[2] * s, p[1], -p[0] * s + p[2] * c], dtype=F32)


def _box_corners(center, size, deg):
    hw, h, hd = F32(size[0]) * F32(0.5), F32(size[1]), F32(size[2]) * F32(0.5)
    local = [(-hw, 0, -hd), (hw, 0, -hd), (-hw, 0, hd), (hw, 0, hd),
             (-hw, h, -hd), (hw, h, -hd), (-hw, h, hd), (hw, h, hd)]
    c = np.asarray(center, dtype=F32)
    return [(_rotate_y(q, deg) + c).astype(F32) for q in local]


def _box5(b: _Builder, center, size, deg, albedo):
    """Five faces (no bottom), cornell_box.hpp:45-90."""
    p = _box_corners(center, size, deg)
    b.quad(p[4], p[5] - p[4], p[6] - p[4], albedo)
    b.quad(p[0], p[1] - p[0], p[4] - p[0], albedo)
    b.quad(p[3], p[2] - p[3], p[7] - p[3], albedo)
    b.quad(p[2], p[0] - p[2], p[6] - p[2], albedo)
    b.quad(p[1], p[3] - p[1], p[5] - p[1], albedo)


WHITE = (0.73, 0.73, 0.73)
RED = (0.65, 0.05, 0.05)
GREEN = (0.12, 0.45, 0.15)
LIGHT_E = (15.0, 15.0, 15.0)


def _walls(b: _Builder, albedos=None):
    """Floor, ceiling, back, left, right (cornell_box.hpp:103-116)."""
    quads = [((-1, 0, -1), (2, 0, 0), (0, 0, 2)),
             ((-1, 2, 1), (2, 0, 0), (0, 0, -2)),
             ((-1, 0, -1), (2, 0, 0), (0, 2, 0)),
             ((-1, 0, 1), (0, 0, -2), (0, 2, 0)),
             ((1, 0, -1), (0, 0, 2), (0, 2, 0))]
    cols = albedos if albedos is not None else [WHITE, WHITE, WHITE, RED, GREEN]
    for (o, u, v), a in zip(quads, cols):
        b.quad(o, u, v, a)


def _ceiling_light(b: _Builder):
    """cornell_box.hpp:119-120."""
    b.quad((-0.25, 1.99, -0.25), (0.5, 0, 0), (0, 0, 0.5), (0.0, 0.0, 0.0), LIGHT_E)


def cornell_ref() -> Mesh:
    b = _Builder()
    _walls(b)
    _ceiling_light(b)
    _box5(b, (-0.35, 0.0, 0.3), (0.5, 1.2, 0.5), 15.0, WHITE)
    _box5(b, (0.35, 0.0, -0.3), (0.55, 0.55, 0.55), -18.0, WHITE)
    return b.mesh(REFERENCE_CAMERA, "cornell_ref")


def _box6_axis_aligned(b: _Builder, center, size, albedos):
    """Six outward-facing faces of an axis-aligned box standing on y = 0."""
    cx, cy, cz = center
    hw, h, hd = size[0] * 0.5, size[1], size[2] * 0.5
    x0, x1, y0, y1, z0, z1 = cx - hw, cx + hw, cy, cy + h, cz - hd, cz + hd
    faces = [((x0, y1, z1), (x1 - x0, 0, 0), (0, 0, z0 - z1)),   # top (+y)
             ((x0, y0, z0), (x1 - x0, 0, 0), (0, 0, z1 - z0)),   # bottom (-y)
             ((x1, y0, z0), (x0 - x1, 0, 0), (0, y1 - y0, 0)),   # back (-z)
             ((x0, y0, z1), (x1 - x0, 0, 0), (0, y1 - y0, 0)),   # front (+z)
             ((x0, y0, z0), (0, 0, z1 - z0), (0, y1 - y0, 0)),   # left (-x)
             ((x1, y0, z1), (0, 0, z0 - z1), (0, y1 - y0, 0))]   # right (+x)
    for (o, u, v), a in zip(faces, albedos):
        b.quad(o, u, v, a)


def cornell36(seed: int = 0) -> Mesh:
    """36 triangles: 5 walls + 2 six-face axis-aligned boxes (at the reference's
    box centres and sizes) + the ceiling light; albedo U[0.1, 0.8] per quad and
    channel from PCG64(seed)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    alb = rng.uniform(0.1, 0.8, size=(17, 3)).astype(F32)
    b = _Builder()
    _walls(b, [tuple(a) for a in alb[:5]])
    _ceiling_light(b)
    _box6_axis_aligned(b, (-0.35, 0.0, 0.3), (0.5, 1.2, 0.5), [tuple(a) for a in alb[5:11]])
    _box6_axis_aligned(b, (0.35, 0.0, -0.3), (0.55, 0.55, 0.55), [tuple(a) for a in alb[11:17]])
    m = b.mesh(REFERENCE_CAMERA, "cornell36")
    assert m.num_triangles == 36
    return m


SOUP_CAMERA = dict(position=(0.0, 0.0, 3.5), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov=40.0)
SOUP_LIGHT_CENTERS = ((-0.5, 1.25, -0.5), (0.5, 1.25, -0.5), (-0.5, 1.25, 0.5), (0.5, 1.25, 0.5))


def soup(num_triangles: int = 1_000_000, seed: int = 1) -> Mesh:
    """Triangle soup: centre c ~ U[-1,1]^3, size s = exp(U[ln 1e-3, ln 5e-2]),
    vertices c + s * d_k with d_k uniform on the unit sphere, albedo U[0,1]^3;
    plus four 0.5 x 0.5 quad lights of radiance 15 at y = 1.25 facing -y
    (8 triangles, appended after the soup)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = int(num_triangles)
    c = rng.uniform(-1.0, 1.0, size=(n, 1, 3))
    s = np.exp(rng.uniform(np.log(1e-3), np.log(5e-2), size=(n, 1, 1)))
    d = rng.normal(size=(n, 3, 3))
    d /= np.linalg.norm(d, axis=2, keepdims=True)
    tri = (c + s * d).astype(F32)
    alb = rng.uniform(0.0, 1.0, size=(n, 3)).astype(F32)
    b = _Builder()
    b.triangles(tri, alb, np.zeros((n, 3), dtype=F32))
    for cx, cy, cz in SOUP_LIGHT_CENTERS:
        b.quad((cx - 0.25, cy, cz - 0.25), (0.5, 0, 0), (0, 0, 0.5), (0.0, 0.0, 0.0), LIGHT_E)
    return b.mesh(SOUP_CAMERA, f"soup{n}")


BATCH_RESOLUTIONS = (128, 256, 512, 1024)


def room_with_soup(num_triangles: int, seed: int) -> Mesh:
    """The 36-triangle Cornell-style room of ``cornell36`` (albedos from this
    seed) filled with num_triangles - 36 soup triangles (centres uniform in
    the room's interior [-0.9, 0.9] x [0.1, 1.9] x [-0.9, 0.9], sizes
    log-uniform in [1e-3, 5e-2], albedo U[0, 1])."""
    room = cornell36(seed)
    extra = max(0, int(num_triangles) - room.num_triangles)
    rng = np.random.Generator(np.random.PCG64(seed + 0x5EED))
    c = rng.uniform((-0.9, 0.1, -0.9), (0.9, 1.9, 0.9), size=(extra, 1, 3))
    s = np.exp(rng.uniform(np.log(1e-3), np.log(5e-2), size=(extra, 1, 1)))
    d = rng.normal(size=(extra, 3, 3))
    d /= np.linalg.norm(d, axis=2, keepdims=True)
    tri = (c + s * d).astype(F32)
    b = _Builder()
    b.v, b.i, b.a, b.e = [room.vertices], [room.indices], [room.albedo], [room.emission]
    b.triangles(tri, rng.uniform(0.0, 1.0, size=(extra, 3)).astype(F32), np.zeros((extra, 3), dtype=F32))
    return b.mesh(REFERENCE_CAMERA, f"room{num_triangles}")


def batch_scenes(num_scenes: int = 64, base_seed: int = 1000, min_tris: float = 1e3, max_tris: float = 5e5,
                 resolutions=BATCH_RESOLUTIONS):
    """Scenes of uneven size for the batched launch: scene i uses seed
    base_seed + i for a triangle count round(exp(U[ln min, ln max])), a
    resolution drawn from ``resolutions`` (square images) and its mesh
    (``room_with_soup``).  Returns [(mesh, width, height, seed)]."""
    out = []
    for i in range(num_scenes):
        seed = base_seed + i
        rng = np.random.Generator(np.random.PCG64(seed))
        t = int(round(np.exp(rng.uniform(np.log(min_tris), np.log(max_tris)))))
        res = int(rng.choice(resolutions))
        out.append((room_with_soup(t, seed), res, res, seed))
    return out


def camera_args(mesh: Mesh) -> dict:
    return dict(mesh.camera)


# ---------------------------------------------------------------- differentiable-rendering scene
DIFF_CAMERA = dict(position=(0.0, 1.1, 2.4), target=(0.0, 0.45, 0.0), up=(0.0, 1.0, 0.0), fov=40.0)


def icosphere(frequency: int, radius: float = 0.5, center=(0.0, 0.5, 0.0)):
    """Geodesic icosphere: every face of the icosahedron split into
    frequency^2 triangles whose grid points are projected onto the sphere.
    Returns (vertices (10 f^2 + 2, 3) float32, indices (20 f^2, 3) uint32),
    triangles wound outwards."""
    f = int(frequency)
    t = (1.0 + np.sqrt(5.0)) / 2.0
    ico_v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                      [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    ico_v /= np.linalg.norm(ico_v, axis=1, keepdims=True)
    ico_f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    key_of = {}
    verts = []

    def vid(p):
        q = p / np.linalg.norm(p)
        k = tuple(np.round(q, 9))
        if k not in key_of:
            key_of[k] = len(verts)
            verts.append(q)
        return key_of[k]

    tris = []
    for a, b, c in ico_f:
        A, B, Cc = ico_v[a], ico_v[b], ico_v[c]
        grid = {}
        for i in range(f + 1):
            for j in range(f + 1 - i):
                grid[(i, j)] = vid(A + (B - A) * (i / f) + (Cc - A) * (j / f))
        for i in range(f):
            for j in range(f - i):
                tris.append((grid[(i, j)], grid[(i + 1, j)], grid[(i, j + 1)]))
                if i + j + 1 < f:
                    tris.append((grid[(i + 1, j)], grid[(i + 1, j + 1)], grid[(i, j + 1)]))
    v = np.asarray(verts) * radius + np.asarray(center)
    idx = np.asarray(tris, dtype=np.uint32)
    # wind outwards
    p0, p1, p2 = v[idx[:, 0]], v[idx[:, 1]], v[idx[:, 2]]
    n = np.cross(p1 - p0, p2 - p0)
    inward = (n * (p0 - np.asarray(center))).sum(axis=1) < 0
    idx[inward] = idx[inward][:, [0, 2, 1]]
    return v.astype(F32), idx


def diff_sphere(frequency: int = 22, stage: str = "target", seed_target: int = 4, seed_init: int = 5) -> Mesh:
    """Differentiable-rendering scene: an icosphere (radius 0.5, resting on the
    ground at the origin; frequency 22 gives 9,680 triangles and 4,842
    vertices) on a 4 x 4 ground quad (albedo 0.6), lit by two quad lights
    facing -y (0.6 x 0.6 at y = 2.0, radiance 12; 0.4 x 0.4 at y = 1.6,
    radiance (10, 8, 6)).
    stage "target": sphere albedos U[0.2, 0.9]^3 per triangle (seed_target);
    stage "init": sphere albedos 0.5 and sphere vertices moved by N(0, 0.01)
    per coordinate (seed_init)."""
    v, idx = icosphere(frequency)
    nt = idx.shape[0]
    if stage == "target":
        alb = np.random.Generator(np.random.PCG64(seed_target)).uniform(0.2, 0.9, size=(nt, 3)).astype(F32)
    elif stage == "init":
        alb = np.full((nt, 3), 0.5, dtype=F32)
        v = (v + np.random.Generator(np.random.PCG64(seed_init)).normal(0.0, 0.01, size=v.shape)).astype(F32)
    else:
        raise ValueError(stage)
    b = _Builder()
    b.v, b.i, b.a, b.e = [v], [idx], [alb], [np.zeros((nt, 3), dtype=F32)]
    b.quad((-2.0, 0.0, 2.0), (4.0, 0.0, 0.0), (0.0, 0.0, -4.0), (0.6, 0.6, 0.6))
    b.quad((-0.9, 2.0, -0.1), (0.6, 0.0, 0.0), (0.0, 0.0, 0.6), (0.0, 0.0, 0.0), (12.0, 12.0, 12.0))
    b.quad((0.6, 1.6, -0.6), (0.4, 0.0, 0.0), (0.0, 0.0, 0.4), (0.0, 0.0, 0.0), (10.0, 8.0, 6.0))
    return b.mesh(DIFF_CAMERA, f"diff_sphere{frequency}_{stage}")


def save_qscene(mesh: Mesh, path) -> None:
    """Writes the host's scene file (little-endian): "QSCN", version 1,
    vertex and triangle counts, camera (position, target, up, fov), then
    vertices, indices, albedo and emission."""
    import struct
    c = mesh.camera
    cam = list(c["position"]) + list(c["target"]) + list(c["up"]) + [c["fov"]]
    with open(path, "wb") as f:
        f.write(b"QSCN")
        f.write(struct.pack("<3I", 1, mesh.vertices.shape[0], mesh.num_triangles))
        f.write(np.asarray(cam, dtype="<f4").tobytes())
        f.write(np.ascontiguousarray(mesh.vertices, dtype="<f4").tobytes())
        f.write(np.ascontiguousarray(mesh.indices, dtype="<u4").tobytes())
        f.write(np.ascontiguousarray(mesh.albedo, dtype="<f4").tobytes())
        f.write(np.ascontiguousarray(mesh.emission, dtype="<f4").tobytes())


# ---------------------------------------------------------------- large heightfield terrain
TERRAIN_CAMERA = dict(position=(0.0, 1.6, 4.6), target=(0.0, 0.0, 0.6), up=(0.0, 1.0, 0.0), fov=45.0)


def _lattice_hash(ix, iz, seed):
    """Lattice values in [0, 1): the PCG hash of the two integer coordinates
    and the octave seed (u32 arithmetic, vectorised)."""
    def pcg(x):
        state = (x * np.uint32(747796405) + np.uint32(2891336453)).astype(np.uint32)
        word = (((state >> ((state >> np.uint32(28)) + np.uint32(4))) ^ state) * np.uint32(277803737)).astype(np.uint32)
        return (word >> np.uint32(22)) ^ word
    with np.errstate(over="ignore"):
        h = pcg(ix.astype(np.uint32) ^ pcg(iz.astype(np.uint32) ^ pcg(np.full_like(ix, seed, dtype=np.uint32))))
    return h.astype(np.float64) * (1.0 / 4294967296.0)


def value_noise(x, z, seed):
    """Value noise: hashed lattice values blended with the smoothstep of the
    fractional coordinates."""
    ix, iz = np.floor(x), np.floor(z)
    fx, fz = x - ix, z - iz
    ix, iz = ix.astype(np.int64), iz.astype(np.int64)
    sx, sz = fx * fx * (3 - 2 * fx), fz * fz * (3 - 2 * fz)
    v00 = _lattice_hash(ix, iz, seed)
    v10 = _lattice_hash(ix + 1, iz, seed)
    v01 = _lattice_hash(ix, iz + 1, seed)
    v11 = _lattice_hash(ix + 1, iz + 1, seed)
    return (v00 * (1 - sx) + v10 * sx) * (1 - sz) + (v01 * (1 - sx) + v11 * sx) * sz


def fbm(x, z, octaves=6, base_frequency=1.5, gain=0.5, seed=7):
    """Fractal sum of value noise in [0, 1): octave k has frequency
    base * 2^k and weight gain^k, normalised by the total weight."""
    total = np.zeros_like(x)
    norm = 0.0
    f, a = base_frequency, 1.0
    for k in range(octaves):
        total += a * value_noise(x * f, z * f, seed + k)
        norm += a
        f *= 2.0
        a *= gain
    return total / norm


def terrain(grid: int = 4096, size: float = 8.0, amplitude: float = 0.2, seed: int = 7) -> Mesh:
    """Heightfield terrain: a grid x grid vertex lattice over [-size/2,
    size/2]^2 (2 (grid-1)^2 triangles; grid 4096 gives 33,538,050), heights
    amplitude * FBM (6 octaves of value noise), per-triangle albedo from the
    noise at the triangle's first vertex; lit by a sun-sized disk (a 16-triangle
    fan of radius 0.25 at distance 50, radiance 4000) and a 40 x 40 sky quad at
    y = 8 (radiance (0.35, 0.45, 0.65))."""
    g = int(grid)
    xs = np.linspace(-size / 2, size / 2, g)
    X, Z = np.meshgrid(xs, xs)                      # row = z, column = x
    n = fbm(X, Z, seed=seed)
    Y = amplitude * n
    verts = np.stack([X, Y, Z], axis=-1).reshape(-1, 3).astype(F32)
    i = np.arange(g - 1, dtype=np.uint32)
    r, c = np.meshgrid(i, i, indexing="ij")
    v00 = (r * g + c).ravel()
    v10 = v00 + 1
    v01 = v00 + g
    v11 = v01 + 1
    tris = np.empty((2 * v00.size, 3), dtype=np.uint32)
    tris[0::2, 0], tris[0::2, 1], tris[0::2, 2] = v00, v01, v11
    tris[1::2, 0], tris[1::2, 1], tris[1::2, 2] = v00, v11, v10
    t = n.reshape(-1)[tris[:, 0]].astype(F32)
    alb = np.stack([0.18 + 0.45 * t, 0.22 + 0.40 * t, 0.12 + 0.25 * t], axis=1).astype(F32)
    b = _Builder()
    b.v, b.i, b.a, b.e = [verts], [tris], [alb], [np.zeros_like(alb)]
    # sun: a fan facing the origin from direction d
    d = np.array([0.45, 0.8, -0.4])
    d /= np.linalg.norm(d)
    centre = 50.0 * d
    u = np.cross(d, [0.0, 0.0, 1.0])
    u /= np.linalg.norm(u)
    w = np.cross(d, u)
    ang = np.linspace(0, 2 * np.pi, 17)
    ring = centre + 0.25 * (np.cos(ang)[:, None] * u + np.sin(ang)[:, None] * w)
    fan = np.stack([np.repeat(centre[None], 16, axis=0), ring[:-1], ring[1:]], axis=1).astype(F32)
    b.triangles(fan, np.zeros((16, 3), dtype=F32), np.full((16, 3), 4000.0, dtype=F32))
    b.quad((-20.0, 8.0, -20.0), (40.0, 0.0, 0.0), (0.0, 0.0, 40.0), (0.0, 0.0, 0.0), (0.35, 0.45, 0.65))
    return b.mesh(TERRAIN_CAMERA, f"terrain{g}")

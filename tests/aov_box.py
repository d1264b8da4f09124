"""An open box of axis-aligned quads whose first-hit AOVs have a closed form.

The camera sits at the origin looking down -z (up +y, fov 40).  Faces, all
with float32-exact coordinates:

- back wall z = -8, x in [-1, 1], y in [-0.75, 0.75]; geometric normal +z
  (towards the camera);
- left wall x = -1, y in [-0.75, 0.75], z in [-8, 0]; geometric normal -x
  (away from the camera, so the facing normal is the flipped one);
- floor y = -0.75, x in [-1, 1], z in [-8, 0]; geometric normal -y (flipped);
- ceiling y = +0.75, emissive (5, 4, 3), same extent; geometric normal -y;
- a small square at z = -0.5, x, y in [-1/32, 1/32]; geometric normal -z
  (flipped), hit at t < 1 so its depth clamps to 0;
- no right wall: rays that leave through x = +1 miss.

``expected_aovs`` computes, without the oracle, what the reference's
pathtrace_frag writes for every pixel and sample (plugin.mm:242-249, :291-307):
the pixel's uv (the fullscreen-triangle uv of plugin.mm:39-44 at the pixel
centre, row 0 at the top, plus the 0.001-uv jitter of plugin.mm:131-132, x
drawn first), the pinhole direction, the nearest face by ray-plane
intersection, then albedo (the emission for an emitter, plugin.mm:243),
normal = facing normal * 0.5 + 0.5 and depth = clamp((t - 1) / 4, 0, 1), all
zero on a miss.  The RNG states come from an independent numpy pcg_hash
(pinned by the known answers of tests/test_oracle_rng.py).
"""
from __future__ import annotations

import math

import numpy as np

from quasi_amd.scenes import Mesh, _Builder

CAMERA = dict(position=(0.0, 0.0, 0.0), target=(0.0, 0.0, -1.0), up=(0.0, 1.0, 0.0), fov=40.0)

BACK = dict(albedo=(0.25, 0.5, 0.75), emission=(0.0, 0.0, 0.0))
LEFT = dict(albedo=(0.125, 0.625, 0.375), emission=(0.0, 0.0, 0.0))
FLOOR = dict(albedo=(0.7, 0.2, 0.1), emission=(0.0, 0.0, 0.0))
CEIL = dict(albedo=(0.3, 0.3, 0.3), emission=(5.0, 4.0, 3.0))
NEAR = dict(albedo=(0.9, 0.1, 0.5), emission=(0.0, 0.0, 0.0))

X, Y, ZB, ZN, S = 1.0, 0.75, 8.0, 0.5, 1.0 / 32.0


def mesh() -> Mesh:
    b = _Builder()
    # quad (o, u, v): geometric normal along u x v
    b.quad((-X, -Y, -ZB), (2 * X, 0, 0), (0, 2 * Y, 0), **BACK)      # u x v = +z
    b.quad((-X, -Y, -ZB), (0, 0, ZB), (0, 2 * Y, 0), **LEFT)         # u x v = -x
    b.quad((-X, -Y, -ZB), (2 * X, 0, 0), (0, 0, ZB), **FLOOR)        # u x v = -y
    b.quad((-X, Y, -ZB), (2 * X, 0, 0), (0, 0, ZB), **CEIL)          # u x v = -y
    b.quad((-S, -S, -ZN), (0, 2 * S, 0), (2 * S, 0, 0), **NEAR)      # u x v = -z
    return b.mesh(CAMERA, "aov_box")


def pcg_hash_np(x):
    x = np.asarray(x, dtype=np.uint64)
    m = np.uint64(0xFFFFFFFF)
    state = (x * np.uint64(747796405) + np.uint64(2891336453)) & m
    shift = (state >> np.uint64(28)) + np.uint64(4)
    word = (((state >> shift) ^ state) * np.uint64(277803737)) & m
    return ((word >> np.uint64(22)) ^ word).astype(np.uint64)


def _to_unit(state):
    return state.astype(np.uint32).astype(np.float32).astype(np.float64) * 2.0 ** -32


def seeds(w, h, sample, rng_mode=0, seed=0):
    y, x = np.mgrid[0:h, 0:w].astype(np.uint64)
    m = np.uint64(0xFFFFFFFF)
    if rng_mode == 0:   # plugin.mm:295
        key = (x + y * np.uint64(1920) + np.uint64(sample) * np.uint64(1920 * 1080)) & m
    else:               # (sample*W*H + y*W + x) ^ pcg_hash(seed)
        key = ((np.uint64(sample) * np.uint64(w * h) + y * np.uint64(w) + x) & m) ^ pcg_hash_np(seed)
    return pcg_hash_np(key)


def pixel_uv(w, h, sample, rng_mode=0, seed=0):
    st = seeds(w, h, sample, rng_mode, seed)
    s1 = pcg_hash_np(st)
    s2 = pcg_hash_np(s1)
    jx, jy = _to_unit(s1), _to_unit(s2)
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    u = (x + 0.5) / w + (jx - 0.5) * 0.001
    v = 1.0 - (y + 0.5) / h + (jy - 0.5) * 0.001
    return u, v


def directions(w, h, sample, rng_mode=0, seed=0):
    """Pinhole at the origin looking down -z: the image plane z = -1 spans
    [-a*t, a*t] x [-t, t] with t = tan(fov/2) and a = w/h."""
    u, v = pixel_uv(w, h, sample, rng_mode, seed)
    t = math.tan(math.radians(CAMERA["fov"]) / 2)
    d = np.stack([(2 * u - 1) * t * (w / h), (2 * v - 1) * t, -np.ones_like(u)], axis=-1)
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


def _face_hits(d):
    """(t, margin) per face for unit directions d (..., 3) from the origin;
    t = inf where the face is not hit, margin = the hit point's distance to the
    face's nearest edge."""
    dx, dy, dz = d[..., 0], d[..., 1], d[..., 2]
    inf = np.inf
    with np.errstate(divide="ignore", invalid="ignore"):
        faces = {}

        def plane(tv, lo_hi):
            p = d * tv[..., None]
            inside = np.ones_like(tv, dtype=bool)
            margin = np.full_like(tv, inf)
            for axis, (lo, hi) in lo_hi.items():
                c = p[..., axis]
                inside &= (c >= lo) & (c <= hi)
                margin = np.minimum(margin, np.minimum(c - lo, hi - c))
            ok = inside & np.isfinite(tv) & (tv > 1e-3)
            return np.where(ok, tv, inf), np.where(ok, margin, inf)

        faces["back"] = plane(np.where(dz < 0, -ZB / dz, inf), {0: (-X, X), 1: (-Y, Y)})
        faces["left"] = plane(np.where(dx < 0, -X / dx, inf), {1: (-Y, Y), 2: (-ZB, 0.0)})
        faces["floor"] = plane(np.where(dy < 0, -Y / dy, inf), {0: (-X, X), 2: (-ZB, 0.0)})
        faces["ceil"] = plane(np.where(dy > 0, Y / dy, inf), {0: (-X, X), 2: (-ZB, 0.0)})
        faces["near"] = plane(np.where(dz < 0, -ZN / dz, inf), {0: (-S, S), 1: (-S, S)})
        # the missing right wall: where the ray leaves the box
        faces["right"] = plane(np.where(dx > 0, X / dx, inf), {1: (-Y, Y), 2: (-ZB, 0.0)})
    return faces


FACING = {   # facing normal (towards the camera) of each face
    "back": (0.0, 0.0, 1.0), "left": (1.0, 0.0, 0.0), "floor": (0.0, 1.0, 0.0),
    "ceil": (0.0, -1.0, 0.0), "near": (0.0, 0.0, 1.0),
}
MATERIAL = {"back": BACK, "left": LEFT, "floor": FLOOR, "ceil": CEIL, "near": NEAR}


def expected_sample(w, h, sample, rng_mode=0, seed=0):
    """Per-pixel first-hit AOVs of one sample: dict of (h, w, 3) arrays
    albedo, normal, depth, emission (the beauty of a depth-1 path), plus
    face names (h, w) and margin (h, w: distance of the hit point to the
    nearest face edge; inf for rays that leave far from any edge)."""
    d = directions(w, h, sample, rng_mode, seed)
    faces = _face_hits(d)
    names = list(faces)
    ts = np.stack([faces[n][0] for n in names])
    ms = np.stack([faces[n][1] for n in names])
    k = np.argmin(ts, axis=0)
    t = np.take_along_axis(ts, k[None], 0)[0]
    margin = np.take_along_axis(ms, k[None], 0)[0]
    which = np.array(names, dtype=object)[k]
    which[~np.isfinite(t)] = "right"
    alb = np.zeros(d.shape)
    nor = np.zeros(d.shape)
    emi = np.zeros(d.shape)
    dep = np.zeros(d.shape)
    for name, mat in MATERIAL.items():
        sel = which == name
        # materials are float32 on the device and in the oracle's scene
        e = np.asarray(mat["emission"], dtype=np.float32).astype(np.float64)
        alb[sel] = e if e.max() > 0.1 else np.asarray(mat["albedo"], dtype=np.float32).astype(np.float64)
        nor[sel] = np.asarray(FACING[name]) * 0.5 + 0.5
        emi[sel] = e
        dep[sel] = np.clip((t[sel] - 1.0) / 4.0, 0.0, 1.0)[:, None]
    # the missing wall: a miss, every AOV zero (its distance to the box edges
    # is the margin that matters)
    return dict(albedo=alb, normal=nor, depth=dep, emission=emi, face=which, margin=margin, t=t)


def expected_sums(w, h, spp, sample_offset=0, rng_mode=0, seed=0):
    out = None
    margin = None
    for s in range(sample_offset, sample_offset + spp):
        e = expected_sample(w, h, s, rng_mode, seed)
        if out is None:
            out = {k: e[k].copy() for k in ("albedo", "normal", "depth", "emission")}
            margin = e["margin"].copy()
        else:
            for k in out:
                out[k] += e[k]
            margin = np.minimum(margin, e["margin"])
    out["margin"] = margin
    return out

"""The quantised 64-byte BVH4 node (quasi_amd/csrc/bvh_quant.hpp) against exact arithmetic.

The traversal kernels test a node's four children against child boxes
quantised to 8 bits per plane; their images stay those of the 128-byte nodes
only if that test never rejects a child box that the ray meets on
[tmin, tmax].  The quantiser and the slab test are compiled for the host from
the kernels' own header (tests/native/quant_shim.cpp), and every check here
is exact (Python fractions over the float inputs):

- every used child's box lies inside its quantised box, on a power-of-two
  grid of at most 255 steps over the union of the children;
- wherever the exact slab test of the original box accepts (the ray's
  float inverse direction, entry/exit planes by its signs, [tmin, tmax]),
  the quantised test accepts: boxes of 1e-7..10 units at offsets up to 1e4,
  flat boxes, identical and nested children, rays aimed at corners, edges and
  faces, axis-parallel rays, origins inside boxes, tmax at the target, and
  rays grazing an edge of planes that sit just off a grid point, far from the
  node's origin (where only the slack on (origin - o) * inv keeps the hit: a
  test with that slack removed or of the wrong sign rejects about a third of
  these hits);
- the quantised test still rejects most exact misses;
- unused slots, child references and the out-of-range flag.
"""
import ctypes
import subprocess
from fractions import Fraction as Fr
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
EMPTY = 0x7FFFFFFF
INF = np.float32(np.inf)


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    out = tmp_path_factory.mktemp("quant") / "libquant_shim.so"
    subprocess.run(["/opt/rocm/bin/hipcc", "-x", "hip", "--offload-arch=gfx950", "-Xclang", "-target-feature",
                    "-Xclang", "-packed-fp32-ops", "-O2", "-std=c++20", "-shared", "-fPIC",
                    "-I", str(ROOT / "quasi_amd" / "csrc"), "-I", str(ROOT / "include"),
                    str(ROOT / "tests" / "native" / "quant_shim.cpp"), "-o", str(out)], check=True)
    lib = ctypes.CDLL(str(out))
    f32p = np.ctypeslib.ndpointer(np.float32, flags="C")
    i32p = np.ctypeslib.ndpointer(np.int32, flags="C")
    lib.qs_quantize.argtypes = [f32p, ctypes.c_uint32, f32p, i32p]
    lib.qs_slab.argtypes = [f32p, f32p, f32p, ctypes.c_float, f32p, ctypes.c_uint32, f32p, f32p]
    return lib


def quantize(lib, nodes):
    nodes = np.ascontiguousarray(nodes, np.float32).reshape(-1, 32)
    out = np.zeros((len(nodes), 16), np.float32)
    ok = np.zeros(len(nodes), np.int32)
    lib.qs_quantize(nodes, len(nodes), out, ok)
    return out, ok


def make_node(boxes, refs):
    """A 128-byte node: boxes[k] = (lo[3], hi[3]) for used slots; refs[k]."""
    n = np.zeros(32, np.float32)
    for k in range(4):
        if refs[k] == EMPTY:
            lo, hi = [INF] * 3, [-INF] * 3
        else:
            lo, hi = boxes[k]
        for a in range(3):
            n[8 * a + k] = lo[a]
            n[8 * a + 4 + k] = hi[a]
    n[24:28] = np.array(refs, np.int32).view(np.float32)
    return n


def random_boxes(rng, kind):
    scale = 10.0 ** rng.uniform(-3, 1)
    offset = 10.0 ** (rng.uniform(3, 4) if kind == "far" else rng.uniform(-2, 1))
    c0 = rng.uniform(-1, 1, 3) * offset
    boxes = []
    for k in range(4):
        c = c0 + rng.uniform(-1, 1, 3) * scale
        h = scale * 10.0 ** rng.uniform(-4 if kind != "tiny" else -7, 0, 3)
        h[rng.random(3) < 0.12] = 0.0
        if kind == "nested" and k > 0:
            c = c0 + rng.uniform(-0.5, 0.5, 3) * scale
            h = scale * 10.0 ** rng.uniform(-6, -3, 3)
        if kind == "nested" and k == 0:
            c, h = c0, np.full(3, scale)
        lo = np.float32(c - h)
        hi = np.float32(c + h)
        boxes.append((np.minimum(lo, hi), np.maximum(lo, hi)))
    if kind == "same":
        boxes = [boxes[0]] * 4
    return boxes


def random_node(rng):
    kind = rng.choice(["near", "far", "tiny", "nested", "same"])
    boxes = random_boxes(rng, str(kind))
    used = int(rng.choice([1, 2, 3, 4], p=[0.05, 0.15, 0.3, 0.5]))
    refs = [int(rng.integers(1, 1 << 20)) if rng.random() < 0.5 else -int(rng.integers(1, 1 << 20))
            for _ in range(4)]
    refs = [refs[k] if k < used else EMPTY for k in range(4)]
    return make_node(boxes, refs), boxes, refs


def safe_rcp(d):
    d = np.float32(d)
    s = np.where(np.abs(d) < np.float32(1e-20), np.copysign(np.float32(1e-20), d), d).astype(np.float32)
    return (np.float32(1.0) / s).astype(np.float32)


def exact_hit(lo, hi, o, inv, tmin, tmax):
    """The slab test of box [lo, hi] in exact arithmetic, with the ray's float inverse direction."""
    tn, tf = Fr(float(tmin)), Fr(float(tmax))
    for a in range(3):
        e, f = (lo[a], hi[a]) if inv[a] >= 0 else (hi[a], lo[a])
        i = Fr(float(inv[a]))
        tn = max(tn, (Fr(float(e)) - Fr(float(o[a]))) * i)
        tf = min(tf, (Fr(float(f)) - Fr(float(o[a]))) * i)
    return tn <= tf


def test_quantised_boxes_contain_their_children(shim):
    rng = np.random.Generator(np.random.PCG64(11))
    nodes, boxes, refs = zip(*(random_node(rng) for _ in range(1500)))
    q, ok = quantize(shim, np.stack(nodes))
    assert ok.all()
    for qi, bx, rf, n in zip(q, boxes, refs, nodes):
        assert np.array_equal(qi[12:16].view(np.int32), np.array(rf, np.int32))
        org = qi[0:3]
        scl = np.array([qi[3], qi[4], qi[5]], np.float32)
        words = qi[6:12].view(np.uint32)   # lo.x hi.x lo.y hi.y lo.z hi.z
        used = [k for k in range(4) if rf[k] != EMPTY]
        for a in range(3):
            m, e = np.frexp(np.float64(scl[a]))
            assert m == 0.5, "scale is a power of two"
            ulo = min(bx[k][0][a] for k in used)
            uhi = max(bx[k][1][a] for k in used)
            assert org[a] == ulo
            assert Fr(255) * Fr(float(scl[a])) >= Fr(float(uhi)) - Fr(float(ulo))
            # not coarser than needed: half the scale would not cover the union (or the exponent is at its floor)
            assert Fr(255) * Fr(float(scl[a])) / 2 < (Fr(float(uhi)) - Fr(float(ulo))) * (1 + Fr(1, 2 ** 19)) \
                or scl[a] == np.float32(2.0 ** -100)
            for k in range(4):
                ql = int((words[2 * a] >> (8 * k)) & 0xFF)
                qh = int((words[2 * a + 1] >> (8 * k)) & 0xFF)
                if k not in used:
                    assert (ql, qh) == (255, 0)
                    continue
                assert Fr(float(org[a])) + ql * Fr(float(scl[a])) <= Fr(float(bx[k][0][a]))
                assert Fr(float(org[a])) + qh * Fr(float(scl[a])) >= Fr(float(bx[k][1][a]))


def rays_at(rng, boxes, refs, per_child):
    """Rays aimed at points of the used children's boxes (corners, edges,
    faces, inside), from near or far origins, some axis-parallel."""
    out = []
    for k in range(4):
        if refs[k] == EMPTY:
            continue
        lo, hi = boxes[k]
        ext = float(np.max(hi - lo)) + 1e-6
        for _ in range(per_child):
            pick = rng.integers(0, 3, 3)
            tgt = np.where(pick == 0, lo, np.where(pick == 1, hi, np.float32((lo + hi) / 2))).astype(np.float32)
            if rng.random() < 0.2:   # origin inside the box
                o = (lo + (hi - lo) * rng.random(3)).astype(np.float32)
                d = rng.normal(size=3).astype(np.float32)
            else:
                o = (tgt + rng.normal(size=3) * ext * 10.0 ** rng.uniform(-1, 3)).astype(np.float32)
                d = (tgt - o).astype(np.float32)
            if rng.random() < 0.15:
                d[rng.integers(0, 3)] = 0.0
            if not np.any(d):
                d[0] = 1.0
            tmin = np.float32(0.0 if rng.random() < 0.5 else 1e-3)
            tmax = np.float32(1.0 if rng.random() < 0.4 else 1e30)
            out.append((o, safe_rcp(d), tmin, tmax))
    return out


@pytest.mark.parametrize("tmin", [0.0, 1e-3])
def test_quantised_slab_test_never_rejects_a_hit(shim, tmin):
    rng = np.random.Generator(np.random.PCG64(5 if tmin == 0.0 else 6))
    nodes, cases = [], []
    for _ in range(1200):
        n, boxes, refs = random_node(rng)
        for (o, inv, _, tmax) in rays_at(rng, boxes, refs, 2):
            nodes.append(n)
            cases.append((boxes, refs, o, inv, tmax))
    q, ok = quantize(shim, np.stack(nodes))
    assert ok.all()
    o = np.stack([c[2] for c in cases]).astype(np.float32)
    inv = np.stack([c[3] for c in cases]).astype(np.float32)
    tmax = np.array([c[4] for c in cases], np.float32)
    tn = np.zeros((len(cases), 4), np.float32)
    tf = np.zeros((len(cases), 4), np.float32)
    shim.qs_slab(np.ascontiguousarray(q), np.ascontiguousarray(o), np.ascontiguousarray(inv), np.float32(tmin),
                 tmax, len(cases), tn, tf)
    accept = tn <= tf
    hits = misses = rejected_misses = 0
    for i, (boxes, refs, oo, ii, tm) in enumerate(cases):
        for k in range(4):
            if refs[k] == EMPTY:
                continue
            if exact_hit(boxes[k][0], boxes[k][1], oo, ii, tmin, tm):
                hits += 1
                assert accept[i, k], (i, k, boxes[k], oo, ii, tm, tn[i, k], tf[i, k])
            else:
                misses += 1
                rejected_misses += int(not accept[i, k])
    print(f"[quant] {len(cases)} rays: {hits} exact hits all accepted; {rejected_misses} of {misses} exact misses "
          f"rejected", flush=True)
    assert hits > 1000 and misses > 1000
    assert rejected_misses >= 0.7 * misses


def grazing_cases(rng, n):
    """A big child spans [-W, W] on x and y; a second child's lo.x and hi.y
    planes lie near 0, just above (lo) and below (hi) a quantisation grid
    point, so their quantised planes keep only 1.25 * 2^-20 of a step of
    margin; rays from near that edge graze it (entering through lo.x at the
    moment they leave through hi.y).  The node's origin is about W away while
    the rays travel well under 1, so the rounding of (origin - o) * inv is far
    above the quantisation margin: only the slack keeps these hits."""
    nodes, cases = [], []
    for _ in range(n):
        w = float(rng.integers(40, 120))
        s = 2.0 ** int(np.ceil(np.log2(2 * w * (1 + 2 ** -20) / 255)))
        lox = np.float32((np.floor(w / s) + 2.0 ** -20 + 2.0 ** -22) * s - w)
        hiy = np.float32((np.ceil(w / s) - 2.0 ** -20 - 2.0 ** -22) * s - w)
        big = (np.float32([-w, -w, -1]), np.float32([w, w, 1]))
        child = (np.float32([lox, hiy - rng.uniform(0.5, 5), -1]), np.float32([lox + rng.uniform(0.5, 5), hiy, 1]))
        d = np.float32([rng.uniform(0.2, 3), rng.uniform(0.2, 3), rng.uniform(-0.1, 0.1)])
        tstar = np.float32(rng.uniform(0.01, 0.5))
        o = np.float32([lox - d[0] * tstar, hiy - d[1] * tstar, rng.uniform(-0.5, 0.5)])
        nodes.append(make_node([big, child], [1, 2, EMPTY, EMPTY]))
        cases.append((child, o, safe_rcp(d)))
    return nodes, cases


def test_quantised_slab_test_keeps_grazing_hits_far_from_the_node_origin(shim):
    rng = np.random.Generator(np.random.PCG64(3))
    nodes, cases = grazing_cases(rng, 4000)
    q, ok = quantize(shim, np.stack(nodes))
    assert ok.all()
    o = np.ascontiguousarray(np.stack([c[1] for c in cases]), np.float32)
    inv = np.ascontiguousarray(np.stack([c[2] for c in cases]), np.float32)
    tmax = np.full(len(cases), 1e30, np.float32)
    tn = np.zeros((len(cases), 4), np.float32)
    tf = np.zeros((len(cases), 4), np.float32)
    shim.qs_slab(np.ascontiguousarray(q), o, inv, np.float32(0.0), tmax, len(cases), tn, tf)
    hits = 0
    for i, (child, oo, ii) in enumerate(cases):
        if exact_hit(child[0], child[1], oo, ii, 0.0, 1e30):
            hits += 1
            assert tn[i, 1] <= tf[i, 1], (i, child, oo, ii, tn[i, 1], tf[i, 1])
    print(f"[quant] grazing rays: {hits} of {len(cases)} exact hits, all accepted", flush=True)
    assert hits > 1000


def test_out_of_range_extent_is_flagged(shim):
    box = (np.float32([-1e30, 0, 0]), np.float32([1e30, 1, 1]))
    n = make_node([box, box, box, box], [1, 2, EMPTY, EMPTY])
    _, ok = quantize(shim, n[None])
    assert ok[0] == 0
    box = (np.float32([-1e17, 0, 0]), np.float32([1e17, 1, 1]))
    _, ok = quantize(shim, make_node([box] * 4, [1, 2, 3, 4])[None])
    assert ok[0] == 1

"""The GPU scene build's traversal layouts (quasi_amd/csrc/bvh_layout.hip)
against the host converter (convert_binary_tree, bvh_build.cpp), which the
CPU tests in test_bvh_layout.py pin: on the same binary tree the pair nodes,
4-wide nodes and triangle order must agree bit for bit, the node counts and
depths exactly and the SAH cost to rounding.  Trees come from the GPU
builder itself (soups, a terrain patch, a Cornell box) and from a random
generator that varies the node numbering, leaf sizes, equal child areas and
inverted (empty) boxes.  Malformed trees must be rejected by the layout's
own index checks, the way the host validator rejects them."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

from quasi_amd import scenes

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
PROBE = ROOT / "tests" / "native" / "liblayoutprobe.so"
BINARY_NODE = np.dtype([("lo", "<f4", 3), ("hi", "<f4", 3), ("left", "<u4"), ("right", "<u4"), ("first", "<u4"),
                        ("count", "<u4")])
LEAF_MAX = 4


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def lib():
    if not PROBE.exists():
        raise RuntimeError(f"{PROBE} is not built (quasi_amd.build.build_test_probes)")
    lib = C.CDLL(str(PROBE))
    P = C.POINTER
    out = [P(C.c_float), P(C.c_float), C.c_uint, P(C.c_float), P(C.c_float), C.c_uint, P(C.c_uint), P(C.c_uint),
           P(C.c_uint), P(C.c_double), C.c_char_p, C.c_uint]
    lib.probe_layout_tree.argtypes = [C.c_void_p, C.c_uint, P(C.c_uint), C.c_uint] + out
    lib.probe_layout_built.argtypes = [P(C.c_float)] * 3 + [C.c_uint] + out
    lib.probe_layout_check.argtypes = [C.c_void_p, C.c_uint, P(C.c_uint), C.c_uint, C.c_char_p, C.c_uint]
    for f in (lib.probe_layout_tree, lib.probe_layout_built, lib.probe_layout_check):
        f.restype = C.c_int
    return lib


def ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def buffers(nodes, ntri):
    """Output arrays for at most `nodes` pair / 4-wide nodes and ntri triangles."""
    cap = max(nodes, 1)
    b = dict(gn=np.zeros(16 * cap, np.float32), hn=np.zeros(16 * cap, np.float32),
             g4=np.zeros(32 * cap, np.float32), h4=np.zeros(32 * cap, np.float32),
             go=np.zeros(ntri, np.uint32), ho=np.zeros(ntri, np.uint32), info=np.zeros(12, np.uint32),
             sah=np.zeros(2, np.float64), why=C.create_string_buffer(256))
    args = [ptr(b["gn"], C.c_float), ptr(b["hn"], C.c_float), cap, ptr(b["g4"], C.c_float), ptr(b["h4"], C.c_float),
            cap, ptr(b["go"], C.c_uint), ptr(b["ho"], C.c_uint), ptr(b["info"], C.c_uint),
            ptr(b["sah"], C.c_double), b["why"], 256]
    return b, args


def compare(b, label):
    info = b["info"]
    g, h = info[0:5], info[5:10]
    say(f"[layout] {label}: GPU nodes/leaves/depth/nodes4/depth4 {list(g)}, host {list(h)}, "
        f"SAH {b['sah'][0]:.9f} / {b['sah'][1]:.9f}")
    assert list(g) == list(h), label
    nn, n4 = int(g[0]), int(g[3])
    assert np.array_equal(b["gn"][:16 * nn].view(np.uint32), b["hn"][:16 * nn].view(np.uint32)), label
    assert np.array_equal(b["g4"][:32 * n4].view(np.uint32), b["h4"][:32 * n4].view(np.uint32)), label
    assert np.array_equal(b["go"], b["ho"]), label
    assert abs(b["sah"][0] - b["sah"][1]) <= 1e-12 * abs(b["sah"][1]), label


def random_tree(n, seed):
    """A binary tree over n triangles: ranges split at random points, leaves of
    1..4 triangles (internal nodes down to 2 triangles), nodes numbered in a
    random expansion order (children always above their parent, left and
    right in either order), random boxes with some inverted (empty) ones and
    some children sharing their sibling's box (equal areas)."""
    rng = np.random.default_rng(seed)
    t = np.zeros(2 * n, dtype=BINARY_NODE)
    lo = rng.uniform(-1, 1, (2 * n, 3)).astype(np.float32)
    ext = rng.uniform(0, 1, (2 * n, 3)).astype(np.float32)
    t["lo"], t["hi"] = lo, lo + ext
    inv = rng.random(2 * n) < 0.02
    t["lo"][inv], t["hi"][inv] = np.float32(np.inf), np.float32(-np.inf)
    frontier = [(0, 0, n)]
    used = 1
    while frontier:
        k = int(rng.integers(len(frontier)))
        frontier[k], frontier[-1] = frontier[-1], frontier[k]
        node, b, c = frontier.pop()
        if node != 0 and c <= LEAF_MAX and (c == 1 or rng.random() < 0.6):
            t[node]["first"], t[node]["count"] = b, c
            continue
        m = int(rng.integers(1, c))
        a, z = (used, used + 1) if rng.random() < 0.5 else (used + 1, used)
        used += 2
        t[node]["left"], t[node]["right"] = a, z
        if rng.random() < 0.1:
            t[z]["lo"], t[z]["hi"] = t[a]["lo"], t[a]["hi"]
        frontier.append((a, b, m))
        frontier.append((z, b + m, c - m))
    return t[:used].copy(), rng.permutation(n).astype(np.uint32)


@pytest.mark.parametrize("n,seed", [(5, 1), (9, 2), (64, 3), (1000, 4), (1000, 5), (60000, 6)])
def test_random_trees(lib, n, seed):
    tree, prims = random_tree(n, seed)
    b, args = buffers(len(tree), n)
    rc = lib.probe_layout_tree(tree.ctypes.data, len(tree), ptr(prims, C.c_uint), n, *args)
    assert rc == 0, (rc, b["why"].value)
    compare(b, f"random tree n={n} seed={seed} ({len(tree)} nodes)")


def corners(mesh):
    v = mesh.vertices.astype(np.float32)
    i = mesh.indices.astype(np.int64)
    return [np.ascontiguousarray(v[i[:, k]]) for k in range(3)]


@pytest.mark.parametrize("name,make", [
    ("cornell36", lambda: scenes.cornell36(2)),
    ("soup 20k", lambda: scenes.soup(20_000, seed=3)),
    ("soup 300k", lambda: scenes.soup(300_000, seed=4)),
    ("terrain 257", lambda: scenes.terrain(257)),
])
def test_trees_of_the_gpu_builder(lib, name, make):
    mesh = make()
    p = corners(mesh)
    n = len(p[0])
    b, args = buffers(2 * n, n)
    rc = lib.probe_layout_built(*[ptr(x, C.c_float) for x in p], n, *args)
    assert rc == 0, (rc, b["why"].value)
    compare(b, f"{name}: {n} triangles, {int(b['info'][10])} tree node slots ({int(b['info'][11])} unused)")


def test_malformed_trees_are_rejected(lib):
    """The same malformed trees the host validator rejects (test_bvh_layout.py);
    every one is caught by the layout's index checks before any node is used."""
    t = np.zeros(5, dtype=BINARY_NODE)
    t["hi"] = 1.0
    t[0]["left"], t[0]["right"] = 1, 2
    t[1]["left"], t[1]["right"] = 3, 4
    t[2]["first"], t[2]["count"] = 0, 3
    t[3]["first"], t[3]["count"] = 3, 2
    t[4]["first"], t[4]["count"] = 5, 2
    prims = np.arange(7, dtype=np.uint32)
    why = C.create_string_buffer(256)
    assert lib.probe_layout_check(t.ctypes.data, 5, ptr(prims, C.c_uint), 7, why, 256) == 0, why.value
    bad = {}
    x = t.copy(); x[1]["left"] = 0; bad["child not above its parent"] = x
    x = t.copy(); x[1]["left"] = 2; bad["node with two parents"] = x
    x = t.copy(); x[0]["right"] = 9; bad["child out of range"] = x
    x = t.copy(); x[4]["count"] = 5; bad["leaf of five triangles"] = x
    x = t.copy(); x[4]["first"] = 6; bad["leaf past the triangles"] = x
    x = t.copy(); x[0]["count"] = 2; bad["leaf root"] = x
    x = t.copy(); x[4]["first"] = 3; bad["triangles counted twice"] = x
    for label, x in bad.items():
        rc = lib.probe_layout_check(x.ctypes.data, 5, ptr(prims, C.c_uint), 7, why, 256)
        say(f"[layout] {label}: rc {rc} ({why.value.decode()})")
        assert rc == 1, label

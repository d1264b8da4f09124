"""The plugin's BVH builder output, checked on the CPU before a kernel walks it.

The builder (quasi_amd/csrc/bvh_build.cpp) is compiled here with g++ through a
tiny test shim.  Checks: every child reference is in range; leaves hold 1..4
triangles and cover each triangle exactly once; each child box contains its
triangles / subtree; the stack depth bound holds; the output does not depend
on the thread count; and a straightforward CPU walk of the pair-node layout
(written here, independently of the kernel) finds the same closest triangle as
brute force.
"""
import ctypes as C
import subprocess

import numpy as np
import pytest

from quasi_amd import scenes
from tests.conftest import ROOT


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    d = tmp_path_factory.mktemp("probe")
    so = d / "libbvhprobe.so"
    subprocess.run(["g++", "-std=c++20", "-O2", "-shared", "-fPIC", "-pthread",
                    "-I", str(ROOT / "quasi_amd" / "csrc"), str(ROOT / "tests" / "native" / "bvh_probe.cpp"),
                    str(ROOT / "quasi_amd" / "csrc" / "bvh_build.cpp"), "-o", str(so)], check=True)
    lib = C.CDLL(str(so))
    P = C.POINTER
    lib.probe_build_cut.restype = C.c_int
    lib.probe_build_cut.argtypes = [P(C.c_float)] * 3 + [C.c_uint, C.c_uint, P(C.c_float), C.c_uint, P(C.c_uint),
                                                          P(C.c_uint), P(C.c_float), C.c_uint, P(C.c_double)]
    lib.probe_validate.restype = C.c_int
    lib.probe_validate.argtypes = [C.c_void_p, C.c_uint, C.c_void_p, C.c_uint]
    lib.probe_build.restype = C.c_int
    lib.probe_build.argtypes = [P(C.c_float)] * 3 + [C.c_uint, C.c_uint, P(C.c_float), C.c_uint, P(C.c_uint),
                                                      P(C.c_uint), P(C.c_float)]
    return lib


def build(probe, mesh, threads=4, with4=False):
    tri = mesh.triangles().astype(np.float32)
    n = tri.shape[0]
    p = [np.ascontiguousarray(tri[:, k, :]) for k in range(3)]
    cap = max(1, 2 * n)
    nodes = np.zeros((cap, 16), dtype=np.float32)
    nodes4 = np.zeros((cap, 32), dtype=np.float32)
    order = np.zeros(max(1, n), dtype=np.uint32)
    info = np.zeros(6, dtype=np.uint32)
    rc = probe.probe_build(*[x.ctypes.data_as(C.POINTER(C.c_float)) for x in p], n, threads,
                           nodes.ctypes.data_as(C.POINTER(C.c_float)), cap,
                           order.ctypes.data_as(C.POINTER(C.c_uint)), info.ctypes.data_as(C.POINTER(C.c_uint)),
                           nodes4.ctypes.data_as(C.POINTER(C.c_float)))
    assert rc == 0
    nn = int(info[0])
    meta = dict(nodes=nn, leaves=int(info[1]), depth=int(info[2]), order=int(info[3]), nodes4=int(info[4]),
                depth4=int(info[5]))
    if with4:
        return nodes[:nn], order[:n], meta, nodes4[:meta["nodes4"]]
    return nodes[:nn], order[:n], meta


def child_refs(nodes):
    ch = nodes[:, 12:14].copy().view(np.int32)
    return ch[:, 0], ch[:, 1]


def boxes(nodes):
    L = (nodes[:, [0, 2, 8]], nodes[:, [1, 3, 9]])
    R = (nodes[:, [4, 6, 10]], nodes[:, [5, 7, 11]])
    return L, R


def check_layout(mesh, nodes, order, info):
    n = mesh.num_triangles
    assert info["order"] == n and sorted(order.tolist()) == list(range(n))
    cl, cr = child_refs(nodes)
    (llo, lhi), (rlo, rhi) = boxes(nodes)
    tri = mesh.triangles()[order]          # BVH order
    tlo, thi = tri.min(1), tri.max(1)
    covered = np.zeros(n, dtype=np.int64)
    parent_count = np.zeros(len(nodes), dtype=np.int64)

    def leaf_range(c):
        v = (~np.int64(c)) & 0xFFFFFFFF
        return int(v >> 2), int(v & 3) + 1

    for i in range(len(nodes)):
        for c, lo, hi in ((cl[i], llo[i], lhi[i]), (cr[i], rlo[i], rhi[i])):
            if c >= 0:
                assert c < len(nodes) and c != 0
                parent_count[c] += 1
            else:
                f, k = leaf_range(c)
                assert 1 <= k <= 4 and f + k <= n
                if not (len(nodes) == 1 and n == 1 and c == cr[i] and i == 0 and cl[i] == cr[i]):
                    covered[f:f + k] += 1
                assert (lo <= tlo[f:f + k]).all() and (hi >= thi[f:f + k]).all()
    if n == 1:
        covered[:] = 1
    assert (covered == 1).all()
    assert parent_count[0] == 0 and (parent_count[1:] == 1).all()
    # subtree containment: a child node's two boxes lie inside the parent's box for it
    for i in range(len(nodes)):
        for c, lo, hi in ((cl[i], llo[i], lhi[i]), (cr[i], rlo[i], rhi[i])):
            if c > 0:
                assert (llo[c] >= lo).all() and (lhi[c] <= hi).all() and (rlo[c] >= lo).all() and (rhi[c] <= hi).all()


def walk_closest(nodes, tri_bvh, order, o, d):
    """Plain recursive-stack walk of the pair nodes with a slab test, then
    closest hit with the larger-input-index tie rule."""
    cl, cr = child_refs(nodes)
    (llo, lhi), (rlo, rhi) = boxes(nodes)
    inv = 1.0 / np.where(np.abs(d) < 1e-20, np.copysign(1e-20, d), d)
    best_t, best = 1e30, -1

    def slab(lo, hi):
        t0, t1 = (lo - o) * inv, (hi - o) * inv
        return max(np.minimum(t0, t1).max(), 1e-3) <= min(np.maximum(t0, t1).min(), best_t) * (1 + 1e-9)

    stack = [0]
    while stack:
        i = stack.pop()
        for c, lo, hi in ((cl[i], llo[i], lhi[i]), (cr[i], rlo[i], rhi[i])):
            if not slab(lo, hi):
                continue
            if c >= 0:
                stack.append(int(c))
                continue
            v = (~np.int64(c)) & 0xFFFFFFFF
            f, k = int(v >> 2), int(v & 3) + 1
            for j in range(f, f + k):
                t = mt(o, d, *tri_bvh[j])
                if t is not None and (t < best_t or (t == best_t and order[j] > best)):
                    best_t, best = t, int(order[j])
    return best, best_t


def mt(o, d, p0, p1, p2):
    e1, e2 = p1 - p0, p2 - p0
    pv = np.cross(d, e2)
    det = e1 @ pv
    if det == 0:
        return None
    s = o - p0
    u = (s @ pv) / det
    if u < 0 or u > 1:
        return None
    q = np.cross(s, e1)
    v = (d @ q) / det
    if v < 0 or u + v > 1:
        return None
    t = (e2 @ q) / det
    return t if 1e-3 <= t <= 1e30 else None


def brute_closest(tris, o, d):
    best_t, best = 1e30, -1
    for j, tr in enumerate(tris):
        t = mt(o, d, *tr)
        if t is not None and (t < best_t or (t == best_t and j > best)):
            best_t, best = t, j
    return best, best_t


@pytest.mark.parametrize("name", ["cornell_ref", "cornell36", "soup3000"])
def test_layout_invariants(probe, name):
    mesh = {"cornell_ref": scenes.cornell_ref, "cornell36": lambda: scenes.cornell36(0),
            "soup3000": lambda: scenes.soup(3000, seed=2)}[name]()
    nodes, order, info = build(probe, mesh)
    check_layout(mesh, nodes, order, info)
    assert info["depth"] <= 32


@pytest.mark.parametrize("ntri", [1, 2, 3, 4, 5, 9])
def test_tiny_scenes(probe, ntri):
    m = scenes.soup(ntri, seed=3)
    m.indices = m.indices[:ntri]
    m.albedo, m.emission = m.albedo[:ntri], m.emission[:ntri]
    nodes, order, info = build(probe, m)
    assert info["nodes"] >= 1
    check_layout(m, nodes, order, info)


def test_thread_count_does_not_change_output(probe):
    m = scenes.soup(200_000, seed=5)
    a = build(probe, m, threads=1)
    b = build(probe, m, threads=8)
    assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32)) and np.array_equal(a[1], b[1])


def test_large_soup_depth_within_stack(probe):
    m = scenes.soup(1_000_000, seed=1)
    nodes, order, info = build(probe, m, threads=8)
    cl, cr = child_refs(nodes)
    internal = np.concatenate([cl[cl >= 0], cr[cr >= 0]])
    assert internal.max() < len(nodes)
    assert info["depth"] <= 64


def test_walk_equals_brute_force(probe):
    m = scenes.soup(2000, seed=6)
    nodes, order, info = build(probe, m)
    tris = m.triangles().astype(np.float64)
    tri_bvh = tris[order]
    rng = np.random.default_rng(1)
    hits = 0
    cen = tris.mean(1)
    for i in range(300):
        o = rng.uniform(-1.3, 1.3, 3)
        d = (cen[rng.integers(len(cen))] - o) if i % 2 else rng.normal(size=3)
        d /= np.linalg.norm(d)
        a = walk_closest(nodes, tri_bvh, order, o, d)
        b = brute_closest(tris, o, d)
        assert a[0] == b[0]
        hits += a[0] >= 0
    assert hits > 10


EMPTY = 0x7FFFFFFF


def check_layout4(mesh, nodes4, order):
    """Every BVH4 reference is in range, every triangle is in exactly one leaf,
    and each child box contains its leaf's triangles / its subtree's boxes."""
    n = mesh.num_triangles
    refs = nodes4[:, 24:28].copy().view(np.int32)
    lo = np.stack([nodes4[:, 0:4], nodes4[:, 8:12], nodes4[:, 16:20]], axis=2)    # (N, 4, 3)
    hi = np.stack([nodes4[:, 4:8], nodes4[:, 12:16], nodes4[:, 20:24]], axis=2)
    tri = mesh.triangles()[order]
    tlo, thi = tri.min(1), tri.max(1)
    covered = np.zeros(n, dtype=np.int64)
    parents = np.zeros(len(nodes4), dtype=np.int64)
    for i in range(len(nodes4)):
        used = refs[i] != EMPTY
        assert used.sum() >= (2 if n > 1 else 1)
        assert (~used[np.argmax(~used):]).all() if (~used).any() else True   # empty slots are trailing
        for k in range(4):
            c = int(refs[i, k])
            if c == EMPTY:
                # an inverted box, which the sign-selected slab test rejects
                assert (lo[i, k] == np.inf).all() and (hi[i, k] == -np.inf).all()
                continue
            if c >= 0:
                assert 0 < c < len(nodes4)
                parents[c] += 1
                assert (lo[c][refs[c] != EMPTY] >= lo[i, k]).all() and (hi[c][refs[c] != EMPTY] <= hi[i, k]).all()
            else:
                v = (~np.int64(c)) & 0xFFFFFFFF
                f, cnt = int(v >> 2), int(v & 3) + 1
                assert f + cnt <= n
                covered[f:f + cnt] += 1
                assert (lo[i, k] <= tlo[f:f + cnt]).all() and (hi[i, k] >= thi[f:f + cnt]).all()
    if n > 1:
        assert (covered == 1).all()
    assert parents[0] == 0 and (parents[1:] == 1).all()
    return refs, lo, hi


def walk4_closest(nodes4, tri_bvh, order, o, d):
    refs = nodes4[:, 24:28].copy().view(np.int32)
    lo = np.stack([nodes4[:, 0:4], nodes4[:, 8:12], nodes4[:, 16:20]], axis=2)
    hi = np.stack([nodes4[:, 4:8], nodes4[:, 12:16], nodes4[:, 20:24]], axis=2)
    inv = 1.0 / np.where(np.abs(d) < 1e-20, np.copysign(1e-20, d), d)
    best_t, best = 1e30, -1
    stack = [0]
    while stack:
        i = stack.pop()
        for k in range(4):
            c = int(refs[i, k])
            if c == EMPTY:
                continue
            t0, t1 = (lo[i, k] - o) * inv, (hi[i, k] - o) * inv
            if max(np.minimum(t0, t1).max(), 1e-3) > min(np.maximum(t0, t1).min(), best_t) * (1 + 1e-9):
                continue
            if c >= 0:
                stack.append(c)
                continue
            v = (~np.int64(c)) & 0xFFFFFFFF
            for j in range(int(v >> 2), int(v >> 2) + int(v & 3) + 1):
                t = mt(o, d, *tri_bvh[j])
                if t is not None and (t < best_t or (t == best_t and order[j] > best)):
                    best_t, best = t, int(order[j])
    return best, best_t


@pytest.mark.parametrize("name", ["cornell_ref", "cornell36", "soup3000"])
def test_layout4_invariants(probe, name):
    mesh = {"cornell_ref": scenes.cornell_ref, "cornell36": lambda: scenes.cornell36(0),
            "soup3000": lambda: scenes.soup(3000, seed=2)}[name]()
    nodes, order, meta, nodes4 = build(probe, mesh, with4=True)
    check_layout4(mesh, nodes4, order)
    assert meta["depth4"] <= meta["depth"]


@pytest.mark.parametrize("ntri", [1, 2, 3, 4, 5, 9])
def test_layout4_tiny_scenes(probe, ntri):
    m = scenes.soup(ntri, seed=3)
    m.indices = m.indices[:ntri]
    m.albedo, m.emission = m.albedo[:ntri], m.emission[:ntri]
    nodes, order, meta, nodes4 = build(probe, m, with4=True)
    assert meta["nodes4"] >= 1
    check_layout4(m, nodes4, order)


def test_walk4_equals_brute_force(probe):
    m = scenes.soup(2000, seed=6)
    nodes, order, meta, nodes4 = build(probe, m, with4=True)
    tris = m.triangles().astype(np.float64)
    tri_bvh = tris[order]
    rng = np.random.default_rng(2)
    cen = tris.mean(1)
    for i in range(300):
        o = rng.uniform(-1.3, 1.3, 3)
        d = (cen[rng.integers(len(cen))] - o) if i % 2 else rng.normal(size=3)
        d /= np.linalg.norm(d)
        assert walk4_closest(nodes4, tri_bvh, order, o, d)[0] == brute_closest(tris, o, d)[0]


def test_layout4_large_soup(probe):
    m = scenes.soup(1_000_000, seed=1)
    nodes, order, meta, nodes4 = build(probe, m, threads=8, with4=True)
    refs = nodes4[:, 24:28].copy().view(np.int32)
    internal = refs[(refs >= 0) & (refs != EMPTY)]
    assert internal.max() < len(nodes4)
    assert meta["nodes4"] < meta["nodes"]
    assert 3 * meta["depth4"] <= 16 + 64     # LDS short stack + global overflow capacity


def build_cut(probe, mesh, cut, threads=4):
    tri = mesh.triangles().astype(np.float32)
    n = tri.shape[0]
    p = [np.ascontiguousarray(tri[:, k, :]) for k in range(3)]
    cap = max(1, 2 * n)
    nodes = np.zeros((cap, 16), dtype=np.float32)
    nodes4 = np.zeros((cap, 32), dtype=np.float32)
    order = np.zeros(max(1, n), dtype=np.uint32)
    info = np.zeros(6, dtype=np.uint32)
    sah = C.c_double(0.0)
    rc = probe.probe_build_cut(*[x.ctypes.data_as(C.POINTER(C.c_float)) for x in p], n, threads,
                               nodes.ctypes.data_as(C.POINTER(C.c_float)), cap,
                               order.ctypes.data_as(C.POINTER(C.c_uint)), info.ctypes.data_as(C.POINTER(C.c_uint)),
                               nodes4.ctypes.data_as(C.POINTER(C.c_float)), cut, C.byref(sah))
    assert rc == 0
    return nodes[:info[0]], nodes4[:info[4]], order[:n], tuple(int(v) for v in info), sah.value


@pytest.mark.parametrize("cut", [1, 3, 6, 0])
def test_parallel_conversion_equals_one_walk(probe, cut):
    """Converting the subtrees below a cut depth on several threads gives the
    layouts of one sequential walk bit for bit (cut 1: almost everything in
    subtrees; 0: the automatic cut, which a 300k soup's tree is large enough
    to use); the SAH cost is the same sum in another association."""
    mesh = scenes.soup(300_000, seed=12)
    seq = build_cut(probe, mesh, 0xFFFFFFFF)
    par = build_cut(probe, mesh, cut)
    for a, b in zip(seq[:3], par[:3]):
        assert a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert seq[3] == par[3]
    assert abs(seq[4] - par[4]) <= 1e-12 * seq[4]
    # and the parallel output is a valid layout
    check_layout4(mesh, par[1], par[2])


BINARY_NODE = np.dtype([("lo", "<f4", 3), ("hi", "<f4", 3), ("left", "<u4"), ("right", "<u4"), ("first", "<u4"),
                        ("count", "<u4")])


def small_tree():
    """Root 0 -> internal 1 and leaf 2; node 1 -> leaves 3 and 4; 7 triangles."""
    t = np.zeros(5, dtype=BINARY_NODE)
    t["hi"] = 1.0
    t[0]["left"], t[0]["right"] = 1, 2
    t[1]["left"], t[1]["right"] = 3, 4
    t[2]["first"], t[2]["count"] = 0, 3
    t[3]["first"], t[3]["count"] = 3, 2
    t[4]["first"], t[4]["count"] = 5, 2
    return t, np.array([6, 5, 4, 3, 2, 1, 0], dtype=np.uint32)


def validate(probe, t, prims, ntri=None):
    return probe.probe_validate(t.ctypes.data, len(t), prims.ctypes.data, len(prims) if ntri is None else ntri)


def test_tree_validation_accepts_a_tree_and_rejects_each_defect(probe):
    assert BINARY_NODE.itemsize == 40
    t, prims = small_tree()
    assert validate(probe, t, prims) == 1
    bad = []
    x = t.copy(); x[1]["left"] = 0; bad.append((x, prims))                    # child index not above the parent
    x = t.copy(); x[1]["left"] = 2; bad.append((x, prims))                    # node 2 has two parents
    x = t.copy(); x[0]["right"] = 9; bad.append((x, prims))                   # child out of range
    x = t.copy(); x[4]["first"] = 4; bad.append((x, prims))                   # position 4 covered twice, 6 never
    x = t.copy(); x[2]["count"] = 5; bad.append((x, prims))                   # leaf larger than kLeafMax
    x = t.copy(); x[4]["count"] = 3; bad.append((x, prims))                   # leaf past the last triangle
    p2 = prims.copy(); p2[0] = 5; bad.append((t, p2))                         # prims not a permutation
    x = np.concatenate([t, t[4:5]]); bad.append((x, prims))                   # node 5 unreachable (no parent)
    for x, p in bad:
        assert validate(probe, x, p) == 0


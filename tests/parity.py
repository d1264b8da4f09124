"""Image and primary-id comparison helpers for GPU-vs-oracle parity tests.

Metric for "mean relative RGB error": the relative L1 error of the image,
sum |gpu - ref| / sum |ref| over all pixels and RGB channels (the bar is
1e-4).  Primary-hit ids must agree on >= 99.99% of pixels, and every
disagreement must be a near tie: the oracle's FP64 primary ray meets both
candidate triangles (with a barycentric slack of 1e-5) at t within 1e-4
relative of each other.
"""
from __future__ import annotations

import numpy as np

MISS = 0xFFFFFFFF


def rel_l1(gpu: np.ndarray, ref: np.ndarray) -> float:
    g = np.asarray(gpu, dtype=np.float64)[..., :3]
    r = np.asarray(ref, dtype=np.float64)[..., :3]
    den = np.abs(r).sum()
    return float(np.abs(g - r).sum() / den) if den > 0 else float(np.abs(g).sum())


def _mt(o, d, p0, p1, p2, slack):
    e1, e2 = p1 - p0, p2 - p0
    pv = np.cross(d, e2)
    det = e1 @ pv
    if det == 0:
        return None
    s = o - p0
    u = (s @ pv) / det
    q = np.cross(s, e1)
    v = (d @ q) / det
    if u < -slack or v < -slack or u + v > 1 + slack:
        return None
    return (e2 @ q) / det


def near_tie(mesh, o, d, tri_a, tri_b, slack=1e-5, rel=1e-4) -> bool:
    if tri_a == MISS or tri_b == MISS:
        # hit vs miss: only a grazing contact with the hit triangle qualifies
        t = tri_a if tri_a != MISS else tri_b
        tri = mesh.triangles()[t].astype(np.float64)
        return _mt(o, d, *tri, slack) is not None and _mt(o, d, *tri, -slack) is None
    ta = _mt(o, d, *mesh.triangles()[tri_a].astype(np.float64), slack)
    tb = _mt(o, d, *mesh.triangles()[tri_b].astype(np.float64), slack)
    if ta is None or tb is None:
        return False
    return abs(ta - tb) <= rel * max(abs(ta), abs(tb))


def check_ids(gpu_ids, ref_ids, mesh, oracle_scene, params, sample=0, min_agree=0.9999):
    agree = gpu_ids == ref_ids
    frac = float(agree.mean())
    bad = np.argwhere(~agree)
    not_ties = []
    for y, x in bad[:200]:
        o, d = oracle_scene.primary_ray(params, int(x), int(y), sample)
        if not near_tie(mesh, o, d, int(gpu_ids[y, x]), int(ref_ids[y, x])):
            not_ties.append((int(x), int(y), int(gpu_ids[y, x]), int(ref_ids[y, x])))
    return frac, bad.shape[0], not_ties

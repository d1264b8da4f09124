"""Writes tests/golden/alg_bytes.json: algorithmic bytes per ray for the
roofline figure, from the oracle's canonical BVH2 traversal (binned SAH, 32
bins, leaf <= 4, near child first, shadow rays stop at the first hit).

  B_ray = 48 + 32 * n_node + 36 * n_tri
    48 B  ray record read (origin, direction, t range: 32 B) + hit write (t, id, u, v: 16 B)
    32 B  one BVH2 child box test (AABB 24 B + child reference 8 B)
    36 B  one triangle test (three float32 vertices)
  plus 32 B per sample for the framebuffer read-modify-write.

n_node and n_tri are the mean box and triangle tests per ray over a fixed
subsample (every 61st pixel; samples and ray types as stated per entry).
Run once, before any GPU measurement:  python tests/golden/make_alg_bytes.py
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import oracle as O  # noqa: E402
from quasi_amd import scenes  # noqa: E402

WORKLOADS = {
    "soup_1m_1080p": dict(mesh=lambda: scenes.soup(1_000_000, seed=1), width=1920, height=1080, spp=4,
                          max_depth=8, rr_start=3, integrator=1, rng_mode=1, seed=0),
    "cornell36_256": dict(mesh=lambda: scenes.cornell36(0), width=256, height=256, spp=16,
                          max_depth=4, rr_start=3, integrator=1, rng_mode=1, seed=0),
}
STRIDE = 61


def entry(rays, aabb, tri):
    n_node = aabb / rays if rays else 0.0
    n_tri = tri / rays if rays else 0.0
    return dict(rays=int(rays), aabb_tests=int(aabb), tri_tests=int(tri), n_node=n_node, n_tri=n_tri,
                bytes_per_ray=48 + 32 * n_node + 36 * n_tri)


def main():
    out = {"formula": "B_ray = 48 + 32*n_node + 36*n_tri; +32 B per sample (framebuffer RMW)",
           "pixel_stride": STRIDE}
    for name, w in WORKLOADS.items():
        mesh = w["mesh"]()
        s = O.Scene(mesh)
        r = s.render(w["width"], w["height"], w["spp"], max_depth=w["max_depth"], rr_start=w["rr_start"],
                     integrator=w["integrator"], rng_mode=w["rng_mode"], seed=w["seed"], pixel_stride=STRIDE,
                     precision=32, threads=0)
        c = r.counts
        e = {t: entry(getattr(c, f"{t}_rays"), getattr(c, f"{t}_aabb"), getattr(c, f"{t}_tri"))
             for t in ("camera", "extension", "shadow")}
        e["all"] = entry(c.camera_rays + c.extension_rays + c.shadow_rays,
                         c.camera_aabb + c.extension_aabb + c.shadow_aabb,
                         c.camera_tri + c.extension_tri + c.shadow_tri)
        e["paths"] = int(c.camera_rays)
        e["sample"] = (f"{w['width']}x{w['height']}, samples 0..{w['spp'] - 1}, every {STRIDE}th pixel, "
                       f"depth {w['max_depth']}, RR from bounce {w['rr_start']}, NEE, hashed seed {w['seed']}")
        e["bvh"] = s.info()
        out[name] = e
        print(name, json.dumps(e["all"]), flush=True)
    (Path(__file__).parent / "alg_bytes.json").write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()

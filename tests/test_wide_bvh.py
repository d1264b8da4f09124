"""What an 8-wide BVH would save, measured on the CPU (tests/native/wide_bvh_steps.cpp).

The soup's BVH2 from the plugin's host builder is collapsed to 4- and 8-wide
trees the way the BVH4 layout is made, and a scalar closest-hit traversal
counts node steps and child-box tests per ray (camera rays and one secondary
ray per camera hit).  The traversal kernel's node step tests every slot of a
node, so its box work per ray is width x steps: going from 4 to 8 children
cuts the steps by about a third but adds about 40% box tests, and the kernel
(quantised nodes) is bound by VALU issue, with the texture addresser's
per-lane loads close behind.
"""
import subprocess

import numpy as np
import pytest

from quasi_amd import scenes

ROOT = __import__("pathlib").Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    out = tmp_path_factory.mktemp("wide") / "wide_bvh_steps"
    subprocess.run(["g++", "-O2", "-std=c++20", "-I", str(ROOT / "quasi_amd" / "csrc"),
                    str(ROOT / "tests" / "native" / "wide_bvh_steps.cpp"),
                    str(ROOT / "quasi_amd" / "csrc" / "bvh_build.cpp"), "-o", str(out), "-pthread"], check=True)
    return out


def test_wider_nodes_fewer_steps_more_box_tests(probe, tmp_path):
    mesh = scenes.soup(200_000, seed=1)
    f = tmp_path / "soup.f32"
    mesh.triangles().astype(np.float32).tofile(f)
    res = subprocess.run([str(probe), str(f), "4000"], capture_output=True, text=True, check=True).stdout
    rows = {}
    for line in res.splitlines():
        t = line.split()
        rows[int(t[1])] = {t[i]: float(t[i + 1]) for i in range(2, len(t), 2)}
    print(res, flush=True)
    s2, s4, s8 = (rows[w]["sec_steps"] for w in (2, 4, 8))
    assert s8 < s4 < s2
    # the same triangles are tested whatever the width (the leaves are the BVH2's)
    assert abs(rows[8]["sec_tris"] - rows[4]["sec_tris"]) <= 0.05 * rows[4]["sec_tris"]
    # slots evaluated per secondary ray grow from 4-wide to 8-wide
    assert 8 * s8 > 1.2 * 4 * s4

"""Pins the oracle's camera and the scene generators.

- Camera: pinhole closed form (a separate formula: tan(fov/2)-scaled image
  plane at unit distance) and the reference's own camera.hpp:36-52, compiled
  from its sources into oracle/_ref and frozen in tests/golden.
- Scene: quasi_amd.scenes.cornell_ref against the reference's own
  make_cornell_box() output (same fixture).
"""
import json
import math

import numpy as np
import pytest

from oracle import oracle as O
from quasi_amd import scenes

CAM = scenes.REFERENCE_CAMERA


def pinhole(u, v, aspect, fov_deg):
    """Camera at +z looking down -z, up +y: the image plane is the square
    [-h*aspect, h*aspect] x [-h, h] at distance 1 with h = tan(fov/2)."""
    h = math.tan(math.radians(fov_deg) / 2)
    d = np.array([(2 * u - 1) * h * aspect, (2 * v - 1) * h, -1.0])
    return d / np.linalg.norm(d)


@pytest.mark.parametrize("u,v", [(0.5, 0.5), (0, 0), (1, 1), (0.25, 0.75), (0.9, 0.1)])
@pytest.mark.parametrize("w,h", [(256, 256), (1920, 1080)])
def test_camera_matches_pinhole_closed_form(u, v, w, h):
    got = O.camera_direction(CAM, w, h, u, v)
    np.testing.assert_allclose(got, pinhole(u, v, w / h, 40.0), atol=1e-12)


def test_camera_known_values():
    np.testing.assert_allclose(O.camera_direction(CAM, 256, 256, 0.5, 0.5), [0, 0, -1], atol=1e-15)
    np.testing.assert_allclose(O.camera_direction(CAM, 256, 256, 0, 0),
                               [-0.32361558, -0.32361558, -0.88912649], atol=5e-8)


def test_camera_matches_reference_camera_hpp(golden):
    ref = json.loads((golden / "ref_cornell_scene.json").read_text())
    cam = ref["camera"]
    assert cam["fov"] == 40 and cam["aspect"] == 1
    np.testing.assert_allclose(cam["position"], CAM["position"])
    for r in ref["rays"]:
        got = O.camera_direction(CAM, 256, 256, r["uv"][0], r["uv"][1])
        # the reference evaluates in float32
        np.testing.assert_allclose(got, r["direction"], atol=2e-7)


def test_oracle_ref_build_agrees_with_fixture(root, golden):
    live = root / "oracle" / "_ref" / "ref_scene.json"
    if not live.exists():
        pytest.skip("reference sources not present; fixture already frozen")
    assert json.loads(live.read_text()) == json.loads((golden / "ref_cornell_scene.json").read_text())


def test_cornell_ref_matches_reference_make_cornell_box(golden):
    ref = json.loads((golden / "ref_cornell_scene.json").read_text())
    m = scenes.cornell_ref()
    quads = ref["quads"]
    assert m.num_triangles == 2 * len(quads) == 32
    tri = m.triangles().astype(np.float64)
    for i, q in enumerate(quads):
        o, u, v = (np.array(q[k], dtype=np.float64) for k in ("origin", "u", "v"))
        np.testing.assert_allclose(tri[2 * i], [o, o + u, o + u + v], atol=2e-7)
        np.testing.assert_allclose(tri[2 * i + 1], [o, o + u + v, o + v], atol=2e-7)
        for t in (2 * i, 2 * i + 1):
            np.testing.assert_allclose(m.albedo[t], q["albedo"], atol=1e-8)
            np.testing.assert_allclose(m.emission[t], q["emission"], atol=1e-8)
    li = ref["light_index"]
    assert set(np.nonzero(m.emission.max(1) > 0.1)[0]) == {2 * li, 2 * li + 1}


def test_cornell_box_corner_closed_forms():
    # tall box: corner 0 = rotate_y((-0.25, 0, -0.25), 15 deg) + (-0.35, 0, 0.3)
    c, s = math.cos(math.radians(15)), math.sin(math.radians(15))
    x, z = -0.25 * c + -0.25 * s, 0.25 * s + -0.25 * c
    m = scenes.cornell_ref()
    # quad 7 (front face of the tall box) starts at corner 0
    np.testing.assert_allclose(m.triangles()[14][0], [x - 0.35, 0, z + 0.3], atol=1e-6)
    np.testing.assert_allclose([x - 0.35, z + 0.3], [-0.65618622, 0.1232233], atol=1e-7)


def test_cornell36_shape_and_albedo_range():
    m = scenes.cornell36(0)
    assert m.num_triangles == 36
    emit = m.emission.max(1) > 0.1
    assert emit.sum() == 2
    a = m.albedo[~emit]
    assert a.min() >= 0.1 and a.max() <= 0.8
    # pairs of triangles share a quad's albedo
    assert np.array_equal(m.albedo[0::2], m.albedo[1::2])
    assert np.array_equal(m.albedo, scenes.cornell36(0).albedo)


def test_soup_statistics():
    m = scenes.soup(20000, seed=1)
    assert m.num_triangles == 20008
    tri = m.triangles()[:20000].astype(np.float64)
    cen = tri.mean(1)
    assert np.abs(cen).max() < 1.06
    edges = np.linalg.norm(tri[:, 1] - tri[:, 0], axis=1)
    assert edges.max() < 0.1 + 1e-6 and np.median(edges) > 1e-3
    assert (m.emission[-8:] == 15).all() and (m.emission[:-8] == 0).all()
    assert np.array_equal(m.vertices, scenes.soup(20000, seed=1).vertices)

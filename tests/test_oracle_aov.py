"""Pins the oracle's first-hit AOVs and its pixel-to-uv mapping with closed forms.

The scene (tests/aov_box.py) is an open box of axis-aligned quads seen from
the origin; every camera ray's first hit, and so every AOV the reference's
pathtrace_frag writes (plugin.mm:242-249, :291-307), follows from ray-plane
intersection.  A misreading of the reference fails here:

- geometric instead of facing normal: three faces have their geometric normal
  pointing away from the camera;
- unclamped depth: the back wall (t >= 8) and the near square (t ~ 0.5) sit on
  both sides of the clamp;
- albedo instead of emission on an emitter: the ceiling;
- no y flip: the ceiling would appear in the bottom rows;
- a missing half-pixel offset or a jitter in pixel rather than uv units: the
  side walls' depth changes with the ray direction;
- misses: the open right side gives zeros in every AOV.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from quasi_amd import scenes
from tests import aov_box

TOL = 1e-9


@pytest.fixture(scope="module")
def box():
    m = aov_box.mesh()
    s = O.Scene(m)
    yield m, s
    s.close()


def test_expected_layout_covers_every_case():
    e = aov_box.expected_sample(48, 32, 0)
    faces = set(e["face"].ravel())
    assert faces == {"back", "left", "floor", "ceil", "near", "right"}
    # the ceiling (emitter) fills the top rows, the floor the bottom rows
    assert (e["face"][0] == "ceil").sum() > 20 and (e["face"][-1] == "floor").sum() > 20
    # depth on both sides of the clamp and inside it
    d = e["depth"][..., 0]
    assert (d == 0).any() and (d == 1).any() and ((d > 0.05) & (d < 0.95)).any()
    t = e["t"][e["face"] == "near"]
    assert t.max() < 1.0


@pytest.mark.parametrize("integrator", [0, 1], ids=["reference", "nee"])
@pytest.mark.parametrize("rng_mode,seed", [(0, 0), (1, 77)], ids=["ref-seed", "hashed"])
def test_oracle_aovs_match_closed_form(box, integrator, rng_mode, seed):
    mesh, s = box
    w, h, spp = 48, 32, 3
    r = s.render(w, h, spp, max_depth=3, rr_start=3, rng_mode=rng_mode, seed=seed, integrator=integrator,
                 precision=64)
    ex = aov_box.expected_sums(w, h, spp, rng_mode=rng_mode, seed=seed)
    assert ex["margin"].min() > 1e-6    # no ray grazes an edge
    for k in ("albedo", "normal", "depth"):
        got = getattr(r, k)
        np.testing.assert_allclose(got[..., :3], ex[k], atol=TOL, rtol=0, err_msg=k)
        assert np.all(got[..., 3] == spp)


@pytest.mark.parametrize("integrator", [0, 1], ids=["reference", "nee"])
def test_oracle_depth_one_beauty_is_emission(box, integrator):
    """With one segment per path the beauty is the emission the camera ray
    meets (plugin.mm:252): the ceiling's radiance, zero elsewhere."""
    mesh, s = box
    w, h = 40, 30
    r = s.render(w, h, 2, max_depth=1, rr_start=3, integrator=integrator, precision=64)
    ex = aov_box.expected_sums(w, h, 2)
    np.testing.assert_allclose(r.beauty[..., :3], ex["emission"], atol=TOL, rtol=0)


def test_oracle_primary_ids_follow_the_faces(box):
    mesh, s = box
    w, h = 48, 32
    r = s.render(w, h, 1, max_depth=1, precision=64)
    e = aov_box.expected_sample(w, h, 0)
    quad = np.where(r.ids == 0xFFFFFFFF, -1, r.ids.astype(np.int64) // 2)
    names = {0: "back", 1: "left", 2: "floor", 3: "ceil", 4: "near", -1: "right"}
    got = np.vectorize(names.get)(quad)
    assert np.array_equal(got, e["face"])
    hit = e["face"] != "right"
    np.testing.assert_allclose(r.tprim[hit], e["t"][hit], rtol=1e-12)


def test_uv_half_pixel_centre_and_y_flip():
    """uv of pixel (x, y) = ((x + 0.5)/W, 1 - (y + 0.5)/H) before the jitter:
    the jitter-free part is recovered exactly from the oracle's primary ray."""
    m = aov_box.mesh()
    s = O.Scene(m)
    w, h = 9, 7
    t = math.tan(math.radians(40.0) / 2)
    for y in range(h):
        for x in range(w):
            _, d = s.primary_ray(dict(width=w, height=h, camera=aov_box.CAMERA), x, y, 0)
            u = (d[0] / -d[2] / (t * w / h) + 1) / 2
            v = (d[1] / -d[2] / t + 1) / 2
            assert abs(u - (x + 0.5) / w) <= 0.0005 + 1e-12
            assert abs(v - (1 - (y + 0.5) / h)) <= 0.0005 + 1e-12
    s.close()


def test_top_row_looks_up_and_centre_pixel_looks_ahead():
    """Reference camera (main.cpp:38-45): row 0 of the image looks up (the ray
    rises, uv.v near 1), the last row down, and the centre pixel of an
    odd-sized image looks along the view direction to within the 0.0005-uv
    jitter."""
    mesh = scenes.cornell_ref()
    s = O.Scene(mesh)
    w, h = 255, 127
    cam = mesh.camera
    view = np.asarray(cam["target"], float) - np.asarray(cam["position"], float)
    view /= np.linalg.norm(view)
    ht = math.tan(math.radians(cam["fov"]) / 2)
    for sample in range(4):
        _, top = s.primary_ray(dict(width=w, height=h), w // 2, 0, sample)
        _, bot = s.primary_ray(dict(width=w, height=h), w // 2, h - 1, sample)
        assert top[1] > 0.3 and bot[1] < -0.3
        # on the image plane at distance 1: (2v - 1) tan(fov/2), v = 1 - 0.5/h
        # up to the jitter's 0.0005 uv
        assert abs(top[1] / -top[2] - ht * (1 - 1 / h)) <= 0.0005 * 2 * ht + 1e-12
        _, c = s.primary_ray(dict(width=w, height=h), w // 2, h // 2, sample)
        # a uv offset of at most 0.0005 moves the direction by at most
        # 0.0005 * 2 * tan(fov/2) * aspect on the image plane at distance 1
        bound = math.atan(0.0005 * 2 * ht * (w / h) * math.sqrt(2))
        ang = math.acos(min(1.0, float(c @ view)))
        assert 0 < ang <= bound
    s.close()

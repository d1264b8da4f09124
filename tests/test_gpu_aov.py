"""GPU first-hit AOVs and pixel mapping against closed forms (no oracle in the loop).

The open box of tests/aov_box.py is rendered through the plugin's C ABI with
every engine and BVH width; the mean albedo, normal and depth AOVs
(plugin.mm:242-249, :301-307), the primary ids and the depth-1 beauty must
equal what ray-plane intersection gives for the same uv (pixel centre, row 0
at the top, 0.001-uv jitter from the pixel's PCG stream).

Bar: every pixel within 2e-6 absolute for albedo / normal (the FP32 mean of
a few equal values) and 1e-5 for depth (FP32 Moeller-Trumbore t on t <= 8);
the primary ids equal the closed form's face on every pixel.  Pixels whose
ray passes within 1e-4 of a face edge in some sample (where FP32 and the
closed form may legitimately pick different faces) are left out; they must be
under 1% of the image.
"""
import numpy as np
import pytest

from quasi_amd import abi
from quasi_amd.plugin import HipRenderer
from tests import aov_box

pytestmark = pytest.mark.gpu

VARIANTS = [
    (abi.Q_ENGINE_MEGAKERNEL, 4, 0),
    (abi.Q_ENGINE_WAVEFRONT, 4, 1),
    (abi.Q_ENGINE_WAVEFRONT, 2, 1),
    (abi.Q_ENGINE_WAVEFRONT, 4, abi.Q_QUEUE_FLAT),
]
VARIANT_IDS = ["megakernel", "wave-bvh4", "wave-bvh2", "wave-flat"]


def render_box(w, h, spp, engine, width, chunk, integrator, rng_mode, seed, max_depth=3):
    mesh = aov_box.mesh()
    c = aov_box.CAMERA
    with HipRenderer(w, h) as r:
        r.set_scene(mesh)
        r.set_camera(c["position"], c["target"], c["up"], c["fov"])
        block = 128 if chunk == abi.Q_QUEUE_FLAT else 256
        r.set_params(spp_per_render=spp, max_depth=max_depth, rr_start=3, seed=seed, rng_mode=rng_mode,
                     integrator=integrator, engine=engine, bvh_width=width, queue_chunk=chunk, block_size=block,
                     samples_per_pass=spp, flags=abi.Q_FLAG_RECORD_IDS, id_sample=0)
        r.render(1)
        return r.readback_aov(), r.primary_ids()


@pytest.mark.parametrize("engine,width,chunk", VARIANTS, ids=VARIANT_IDS)
@pytest.mark.parametrize("integrator,rng_mode,seed", [(0, 0, 0), (1, 1, 77)], ids=["reference", "nee-hashed"])
def test_gpu_aovs_match_closed_form(engine, width, chunk, integrator, rng_mode, seed):
    w, h, spp = 96, 64, 3
    print(f"[aov-box] engine={engine} width={width} chunk={chunk} integrator={integrator} {w}x{h}x{spp}",
          flush=True)
    aov, ids = render_box(w, h, spp, engine, width, chunk, integrator, rng_mode, seed)
    ex = aov_box.expected_sums(w, h, spp, rng_mode=rng_mode, seed=seed)
    clear = ex["margin"] > 1e-4
    print(f"[aov-box] {int((~clear).sum())} of {w * h} pixels graze an edge", flush=True)
    assert clear.mean() > 0.99
    for k, tol in (("albedo", 2e-6), ("normal", 2e-6), ("depth", 1e-5)):
        got = aov[k][..., :3].astype(np.float64)
        err = np.abs(got - ex[k] / spp)[clear].max()
        print(f"[aov-box] {k}: max abs error {err:.3e}", flush=True)
        assert err <= tol, (k, err)
        assert np.all(aov[k][..., 3] == 1.0)
    e0 = aov_box.expected_sample(w, h, 0, rng_mode, seed)
    quad = np.where(ids == 0xFFFFFFFF, -1, ids.astype(np.int64) // 2)
    names = {0: "back", 1: "left", 2: "floor", 3: "ceil", 4: "near", -1: "right"}
    sel = e0["margin"] > 1e-4
    assert np.array_equal(np.vectorize(names.get)(quad)[sel], e0["face"][sel])


@pytest.mark.parametrize("engine,width,chunk", VARIANTS, ids=VARIANT_IDS)
def test_gpu_depth_one_beauty_is_emission(engine, width, chunk):
    w, h, spp = 64, 48, 2
    print(f"[aov-box depth 1] engine={engine} width={width} chunk={chunk}", flush=True)
    aov, _ = render_box(w, h, spp, engine, width, chunk, 0, 0, 0, max_depth=1)
    ex = aov_box.expected_sums(w, h, spp)
    err = np.abs(aov["beauty"][..., :3].astype(np.float64) - ex["emission"] / spp).max()
    assert err <= 1e-5, err

"""GPU binned-SAH BVH build (Q_BVH_BUILD_GPU) against the host build.

Closest hits do not depend on the tree (the tie rule picks the larger input
index whatever the traversal order), so a valid GPU-built BVH must render the
same image and primary ids, bit for bit, as the host-built one, for every
engine and BVH width.  Its SAH cost (same bins, same leaf rule) is within 3%
of the host tree's.  The plugin validates the downloaded tree (every node
reached once, leaves of 1..4 triangles covering every position once, a
permutation of the triangles) before using it.
"""
import numpy as np
import pytest

from quasi_amd import abi, scenes
from quasi_amd.plugin import HipRenderer, make_camera

pytestmark = pytest.mark.gpu


def render(mesh, builder, w, h, engine, width, spp=4):
    with HipRenderer(w, h) as r:
        r.set_params(bvh_builder=builder)
        r.set_scene(mesh)
        st = r.stats()
        c = mesh.camera
        r.camera = make_camera(c["position"], c["target"], c["up"], c["fov"])
        r.set_params(spp_per_render=spp, max_depth=6, rr_start=3, rng_mode=1, integrator=1, seed=5,
                     engine=engine, bvh_width=width, flags=abi.Q_FLAG_RECORD_IDS, id_sample=0,
                     bvh_builder=builder)
        r.render(1)
        return r.readback_aov(), r.primary_ids(), st


def degenerate_mesh():
    """Many copies of a few triangles (coincident centroids) plus a floor."""
    base = scenes.soup(40, seed=3)
    v = np.concatenate([base.vertices] * 200)
    i = np.concatenate([base.indices + k * base.vertices.shape[0] for k in range(200)]).astype(np.uint32)
    a = np.concatenate([base.albedo] * 200)
    e = np.concatenate([base.emission] * 200)
    return scenes.Mesh(v.astype(np.float32), i, a.astype(np.float32), e.astype(np.float32), dict(base.camera), "dup")


CASES = [("soup200k", lambda: scenes.soup(200_000, seed=4)),
         ("cornell36", lambda: scenes.cornell36(0)),
         ("room20k", lambda: scenes.room_with_soup(20_000, 9)),
         ("duplicates", degenerate_mesh)]


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_gpu_build_renders_like_host_build(name, make):
    mesh = make()
    w, h = 96, 64
    print(f"[gpu-bvh] {name}: {mesh.num_triangles} triangles", flush=True)
    for engine, width in ((abi.Q_ENGINE_WAVEFRONT, 4), (abi.Q_ENGINE_WAVEFRONT, 2), (abi.Q_ENGINE_MEGAKERNEL, 2)):
        a_host, id_host, st_host = render(mesh, abi.Q_BVH_BUILD_HOST, w, h, engine, width)
        a_gpu, id_gpu, st_gpu = render(mesh, abi.Q_BVH_BUILD_GPU, w, h, engine, width)
        assert st_host.bvh_builder == abi.Q_BVH_BUILD_HOST
        assert st_gpu.bvh_builder == abi.Q_BVH_BUILD_GPU, "the GPU build fell back to the host"
        for k in ("beauty", "albedo", "normal", "depth"):
            assert np.array_equal(a_gpu[k], a_host[k]), (name, engine, width, k)
        assert np.array_equal(id_gpu, id_host)
    print(f"[gpu-bvh] {name}: SAH host {st_host.bvh_sah_cost:.3f} gpu {st_gpu.bvh_sah_cost:.3f}, "
          f"nodes {st_host.bvh_nodes}/{st_gpu.bvh_nodes}, depth {st_host.bvh_depth}/{st_gpu.bvh_depth}, "
          f"gpu build {st_gpu.bvh_device_ms:.2f} ms (set_scene {st_gpu.bvh_build_ms:.0f} ms vs "
          f"{st_host.bvh_build_ms:.0f} ms)", flush=True)
    assert st_gpu.bvh_sah_cost <= 1.03 * st_host.bvh_sah_cost
    # both builders make the same binned-SAH decisions on these scenes: the same tree
    assert st_gpu.bvh_nodes == st_host.bvh_nodes and st_gpu.bvh_depth == st_host.bvh_depth
    assert abs(st_gpu.bvh_sah_cost - st_host.bvh_sah_cost) <= 1e-9 * st_host.bvh_sah_cost


def test_auto_builder_uses_gpu_for_large_scenes():
    mesh = scenes.soup(1_100_000, seed=1)
    with HipRenderer(32, 32) as r:
        r.set_scene(mesh)
        st = r.stats()
    print(f"[gpu-bvh] auto 1.1M: builder {st.bvh_builder}, {st.bvh_device_ms:.1f} ms device, "
          f"{st.bvh_build_ms:.0f} ms set_scene, SAH {st.bvh_sah_cost:.2f}", flush=True)
    assert st.bvh_builder == abi.Q_BVH_BUILD_GPU

"""The large workload: a 4096 x 4096 heightfield (33,538,050 terrain
triangles, FBM heights) lit by a sun disk and a sky quad, 2048 x 2048.

Size-independent properties checked at full size:
- the GPU-built BVH renders the same image and primary ids, bit for bit, as
  the host-built BVH (512 x 512, 1 spp);
- Euler's identity for the adjoint at 2048 x 2048, 16 spp, depth 6 without
  Russian roulette: sum(albedo * dl/dalbedo) = sum_k (k-1) (l_k - l_{k-1}),
  l_k the loss of GPU forward renders of depth k (each NEE term found at
  bounce b carries b+1 albedo factors), to 1e-4 relative;
- the 1M-parameter albedo grid (1024 x 1024 blocks of 4 x 4 cells, 32
  triangles each) collects every terrain triangle's gradient once;
- with block-constant albedo (the grid as the scene's parameters), the
  device's FP64 reduction onto the grid satisfies Euler's identity on the grid
  itself, and passes bounded to 2^20 paths (single samples of 512-row bands:
  per-tile path recomputation) give the same grid gradient as the default
  passes.
"""
import time

import numpy as np
import pytest

from quasi_amd import abi, scenes
from quasi_amd.plugin import HipRenderer

pytestmark = pytest.mark.gpu

GRID = 4096
NT_TERRAIN = 2 * (GRID - 1) ** 2


@pytest.fixture(scope="module")
def terrain():
    t0 = time.time()
    m = scenes.terrain(GRID)
    print(f"[terrain] {m.num_triangles} triangles generated in {time.time() - t0:.1f} s", flush=True)
    return m


def cam(r, m):
    c = m.camera
    r.set_camera(c["position"], c["target"], c["up"], c["fov"])


def test_gpu_and_host_bvh_render_alike(terrain):
    out = {}
    for builder in (abi.Q_BVH_BUILD_GPU, abi.Q_BVH_BUILD_HOST):
        with HipRenderer(512, 512) as r:
            r.set_params(bvh_builder=builder)
            t0 = time.time()
            r.set_scene(terrain)
            st = r.stats()
            print(f"[terrain] builder {builder}: set_scene {time.time() - t0:.1f} s (build+upload "
                  f"{st.bvh_build_ms:.0f} ms, device {st.bvh_device_ms:.0f} ms), SAH {st.bvh_sah_cost:.2f}, "
                  f"nodes {st.bvh_nodes}, depth {st.bvh_depth}", flush=True)
            for msg in r.messages:
                if msg.startswith("scene:"):
                    print(f"[terrain] {msg}", flush=True)
            assert st.bvh_builder == builder
            cam(r, terrain)
            r.set_params(spp_per_render=1, max_depth=6, rr_start=3, rng_mode=1, integrator=1, seed=2,
                         engine=abi.Q_ENGINE_WAVEFRONT, bvh_width=4, flags=abi.Q_FLAG_RECORD_IDS, bvh_builder=builder)
            r.render(1)
            out[builder] = (r.readback(), r.primary_ids(), st.bvh_sah_cost)
    (ia, ida, ca), (ib, idb, cb) = out[abi.Q_BVH_BUILD_GPU], out[abi.Q_BVH_BUILD_HOST]
    assert np.array_equal(ia, ib) and np.array_equal(ida, idb)
    assert ca <= 1.03 * cb


def test_adjoint_euler_identity_full_size(terrain):
    w = h = 2048
    spp, D = 16, 6
    g = np.random.Generator(np.random.PCG64(1)).normal(size=(h, w, 4)).astype(np.float32)
    g[..., 3] = 0.0
    with HipRenderer(w, h) as r:
        r.set_scene(terrain)
        cam(r, terrain)
        losses = []
        for k in range(1, D + 1):
            r.set_params(spp_per_render=spp, max_depth=k, rr_start=D, rng_mode=1, integrator=1, seed=0,
                         engine=abi.Q_ENGINE_WAVEFRONT, bvh_width=4, flags=abi.Q_FLAG_TIME_KERNELS)
            r.reset_stats()
            r.render(1)
            img = r.readback().astype(np.float64)
            losses.append(float((img[..., :3] * g[..., :3]).sum()))
            if k == D:
                st = r.stats()
                rays = st.rays_camera + st.rays_extension + st.rays_shadow
                print(f"[terrain] forward {w}x{h} {spp} spp depth {D}: {rays} rays, "
                      f"{st.render_ms:.1f} ms kernels", flush=True)
        t0 = time.time()
        ga, gv, res = r.adjoint(g, spp)
        print(f"[terrain] adjoint: {res.kernel_ms:.1f} ms device, {time.time() - t0:.2f} s wall", flush=True)
    alb = terrain.albedo.astype(np.float64)
    lhs = float((alb * ga).sum())
    rhs = sum((k - 1) * (losses[k - 1] - losses[k - 2]) for k in range(2, D + 1))
    print(f"[terrain] Euler: sum(albedo * grad) {lhs:.6f}, from forward renders {rhs:.6f}", flush=True)
    assert abs(lhs - rhs) <= 1e-4 * abs(rhs)
    assert np.isfinite(gv).all() and np.abs(gv).sum() > 0
    # 1024 x 1024 parameter grid over the (GRID-1)^2 cells, 2 triangles per cell
    cells = np.arange(NT_TERRAIN) // 2
    row, col = cells // (GRID - 1), cells % (GRID - 1)
    group = np.minimum(row // 4, 1023) * 1024 + np.minimum(col // 4, 1023)
    gg = np.stack([np.bincount(group, weights=ga[:NT_TERRAIN, c], minlength=1024 * 1024) for c in range(3)], axis=1)
    assert gg.shape == (1024 * 1024, 3)
    assert np.allclose(gg.sum(axis=0), ga[:NT_TERRAIN].sum(axis=0), rtol=1e-9, atol=1e-12)


def grid_map():
    """Albedo parameter of every triangle: the 1024 x 1024 grid of 4 x 4-cell
    blocks (the last block of a row or column also takes the 3 leftover
    cells); the sun and sky triangles have none."""
    cells = np.arange(NT_TERRAIN, dtype=np.int64) // 2
    row, col = cells // (GRID - 1), cells % (GRID - 1)
    return (np.minimum(row // 4, 1023) * 1024 + np.minimum(col // 4, 1023)).astype(np.uint32)


def test_adjoint_parameter_grid_full_size(terrain):
    w = h = 2048
    spp, D = 16, 6
    npar = 1024 * 1024
    group = grid_map()
    # block-constant albedo: the block mean of the noise albedo
    alb = terrain.albedo.astype(np.float64)
    theta = np.stack([np.bincount(group, weights=alb[:NT_TERRAIN, c], minlength=npar) for c in range(3)], axis=1)
    theta /= np.bincount(group, minlength=npar)[:, None]
    mesh = scenes.Mesh(terrain.vertices, terrain.indices, terrain.albedo.copy(), terrain.emission, terrain.camera,
                       "terrain-grid")
    mesh.albedo[:NT_TERRAIN] = theta[group].astype(np.float32)
    theta32 = theta.astype(np.float32).astype(np.float64)
    pmap = np.full(mesh.num_triangles, abi.Q_ADJOINT_NO_PARAM, dtype=np.uint32)
    pmap[:NT_TERRAIN] = group
    g = np.random.Generator(np.random.PCG64(3)).normal(size=(h, w, 4)).astype(np.float32)
    g[..., 3] = 0.0
    with HipRenderer(w, h) as r:
        r.set_scene(mesh)
        cam(r, mesh)
        losses = []
        for k in range(1, D + 1):
            r.set_params(spp_per_render=spp, max_depth=k, rr_start=D, rng_mode=1, integrator=1, seed=5,
                         engine=abi.Q_ENGINE_WAVEFRONT, bvh_width=4)
            r.render(1)
            img = r.readback().astype(np.float64)
            losses.append(float((img[..., :3] * g[..., :3]).sum()))
        t0 = time.time()
        gp, gv, res = r.adjoint(g, spp, albedo_param=pmap, num_albedo_params=npar)
        print(f"[terrain-grid] adjoint onto {npar} parameters: {res.kernel_ms:.1f} ms device, "
              f"{time.time() - t0:.2f} s wall, {res.passes} passes, {res.working_bytes / 2**30:.2f} GiB", flush=True)
        t0 = time.time()
        gb, vb, rb = r.adjoint(g, spp, albedo_param=pmap, num_albedo_params=npar, max_pass_paths=1 << 20)
        print(f"[terrain-grid] 2^20-path passes: {rb.passes} passes in {rb.bands} bands, {rb.kernel_ms:.1f} ms "
              f"device, {rb.working_bytes / 2**30:.2f} GiB", flush=True)
        ga, _, _ = r.adjoint(g, spp)
    host = np.stack([np.bincount(group, weights=ga[:NT_TERRAIN, c], minlength=npar) for c in range(3)], axis=1)
    del ga
    err_map = float(np.abs(gp - host).sum() / np.abs(host).sum())
    print(f"[terrain-grid] device reduction onto the grid vs per-triangle gradients summed on the host: "
          f"rel-L1 {err_map:.2e}", flush=True)
    assert err_map <= 1e-6
    assert (rb.bands, rb.passes) == (4, 4 * spp)
    assert rb.working_bytes < res.working_bytes
    lhs = float((theta32 * gp).sum())
    rhs = sum((k - 1) * (losses[k - 1] - losses[k - 2]) for k in range(2, D + 1))
    print(f"[terrain-grid] Euler on the grid: sum(theta * grad) {lhs:.6f}, from forward renders {rhs:.6f}", flush=True)
    assert abs(lhs - rhs) <= 1e-4 * abs(rhs)
    err = float(np.abs(gb - gp).sum() / np.abs(gp).sum())
    print(f"[terrain-grid] banded vs default passes: rel-L1 {err:.2e}", flush=True)
    assert err <= 1e-5 and float(np.abs(vb - gv).sum() / np.abs(gv).sum()) <= 1e-5

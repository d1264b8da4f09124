"""The adjoint behind a thin torch.autograd.Function (test harness only: the
library itself has no torch dependency).

forward(albedo, vertices) sets the scene and renders the image through the
plugin; backward(dL/dI) calls Q_plugin_adjoint for the same samples.  Checked:
autograd's gradients equal a direct adjoint call, and Adam steps on the
sphere's albedo lowers an L2 image loss against a target render.
"""
from dataclasses import replace

import numpy as np
import pytest

from quasi_amd import abi, scenes
from quasi_amd.plugin import HipRenderer

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


class RenderFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, albedo, vertices, r, mesh, spp):
        m = replace(mesh, albedo=albedo.detach().cpu().numpy().astype(np.float32),
                    vertices=vertices.detach().cpu().numpy().astype(np.float32))
        r.set_scene(m)
        c = m.camera
        r.set_camera(c["position"], c["target"], c["up"], c["fov"])
        r.set_params(spp_per_render=spp, max_depth=4, rr_start=3, rng_mode=1, integrator=1, seed=1,
                     engine=abi.Q_ENGINE_WAVEFRONT, bvh_width=4, block_size=256, queue_chunk=64)
        r.render(1)
        ctx.r, ctx.spp = r, spp
        return torch.from_numpy(r.readback_aov()["beauty"][..., :3].copy())

    @staticmethod
    def backward(ctx, gout):
        g = gout.detach().cpu().numpy()
        dimg = np.zeros(g.shape[:2] + (4,), dtype=np.float32)
        dimg[..., :3] = g
        ga, gv, _ = ctx.r.adjoint(dimg, ctx.spp)
        return torch.from_numpy(ga).float(), torch.from_numpy(gv).float(), None, None, None


def test_autograd_matches_direct_adjoint():
    w, h, spp = 64, 64, 8
    init = scenes.diff_sphere(8, "init")
    target = scenes.diff_sphere(8, "target")
    with HipRenderer(w, h) as r:
        T = RenderFn.apply(torch.from_numpy(target.albedo), torch.from_numpy(target.vertices), r, target, spp)
        alb = torch.from_numpy(init.albedo.copy()).requires_grad_(True)
        ver = torch.from_numpy(init.vertices.copy()).requires_grad_(True)
        img = RenderFn.apply(alb, ver, r, init, spp)
        loss = ((img - T.detach()) ** 2).sum()
        loss.backward()
        dimg = np.zeros((h, w, 4), dtype=np.float32)
        dimg[..., :3] = 2.0 * (img.detach().numpy() - T.numpy())
        ga, gv, _ = r.adjoint(dimg, spp)
    assert np.allclose(alb.grad.numpy(), ga.astype(np.float32), rtol=1e-5, atol=1e-8)
    assert np.allclose(ver.grad.numpy(), gv.astype(np.float32), rtol=1e-5, atol=1e-8)


def test_albedo_descent_lowers_the_loss():
    w, h, spp = 64, 64, 16
    init = scenes.diff_sphere(8, "init")
    target = scenes.diff_sphere(8, "target")
    ver = torch.from_numpy(target.vertices.copy())
    with HipRenderer(w, h) as r:
        T = RenderFn.apply(torch.from_numpy(target.albedo), ver, r, target, spp).detach()
        alb = torch.from_numpy(init.albedo.copy()).requires_grad_(True)
        opt = torch.optim.Adam([alb], lr=0.04)
        losses = []
        for _ in range(15):
            opt.zero_grad()
            loss = ((RenderFn.apply(alb, ver, r, init, spp) - T) ** 2).sum()
            loss.backward()
            opt.step()
            with torch.no_grad():
                alb.clamp_(0.0, 1.0)
            losses.append(float(loss))
    print("losses", [round(x, 4) for x in losses], flush=True)
    assert losses[-1] < 0.5 * losses[0]

"""Differentiable density fitting over a ring of views.

`DensityFit(fw)` renders the ground truth of a `configs.FitWorkload` from
every view (targets, per-pixel means), and fits the FOG brick values of
`fw.init` to them with an L2 loss on the images and Adam, differentiating
the views' renders with the path-replay adjoint (`autograd.render_views`:
one event reduction for all views of a step). Brick
values are clamped to [0, majorant] after each step, as the adjoint keeps
the majorants fixed.
"""
import numpy as np

from . import autograd as AG
from .kernel import RenderKernel
from .sampler import HosekSky, MiePhase, SunLight
from .camera import sun_direction


def _kernel(scene, viewi, proji, w, h, scattevs, sigma_t, albedo=0.99):
    k = RenderKernel()
    k.Initialize(scene, viewi, proji, MiePhase(), HosekSky(3.0, 0.1), None, scattevs, sigma_t * albedo,
                 sigma_t * (1.0 - albedo), w, h, w, h, 0, [SunLight(sun_direction(30.0, 0.0))])
    return k


class DensityFit:
    def __init__(self, fw, target_spp=None, target_seed=1 << 20):
        import torch
        self.fw = fw
        self.torch = torch
        self.kernels = [_kernel(fw.init, vi, pi, fw.w, fw.h, fw.scattevs, fw.sigma_t) for vi, pi in fw.views]
        gt_k = [_kernel(fw.gt.scene, vi, pi, fw.w, fw.h, fw.scattevs, fw.sigma_t) for vi, pi in fw.views]
        spp = target_spp or 4 * fw.spp
        n = fw.w * fw.h * 4
        self.targets = []
        for i, k in enumerate(gt_k):
            ol, oe = torch.empty(n, device="cuda"), torch.empty(n, device="cuda")
            k.render_device(0, 0, fw.w, fw.h, spp, target_seed + i, 0, ol, oe)
            torch.cuda.synchronize()
            img = (ol + oe).view(fw.h, fw.w, 4)[..., :3] / spp
            self.targets.append(img.clone())
            k.Destroy()
        # per-value bound: the majorant of the leaf that owns each brick
        vol = fw.init.ob["volx"][:, 1]
        q = fw.init.ob["qval"][:, 1]
        lim = np.zeros(fw.init.leafx[1], np.float32)
        m = vol != 0xFFFFFFFF
        lim[vol[m]] = q[m]
        self.qmax = torch.tensor(np.repeat(lim, fw.init.lvoxc3), device="cuda")

    def loss(self, rho, seed, spp=None, views=None):
        """Sum over views of the squared error of the per-pixel mean RGB."""
        spp = spp or self.fw.spp
        idx = list(range(len(self.kernels)) if views is None else views)
        imgs = AG.render_views([self.kernels[i] for i in idx], rho, spp, [seed * 1009 + i for i in idx])
        total = 0.0
        for i, (il, ie) in zip(idx, imgs):
            img = (il + ie)[..., :3] / spp
            total = total + ((img - self.targets[i]) ** 2).sum()
        return total

    def run(self, iters, lr=0.05, seed=0, rho=None, log=None):
        torch = self.torch
        if rho is None:
            rho = torch.tensor(self.fw.init.pvol[1], device="cuda")
        rho = rho.clone().requires_grad_(True)
        opt = torch.optim.Adam([rho], lr=lr)
        losses = []
        for it in range(iters):
            opt.zero_grad()
            loss = self.loss(rho, seed + it)
            loss.backward()
            opt.step()
            with torch.no_grad():
                rho.clamp_(min=0.0)
                rho.copy_(torch.minimum(rho, self.qmax))
            losses.append(float(loss.detach()))
            if log:
                log(it, losses[-1])
        return rho.detach(), losses

    def destroy(self):
        for k in self.kernels:
            k.Destroy()
        self.kernels = []

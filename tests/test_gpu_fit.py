"""Density fitting (the C3 workload) on the GPU.

1. Per-voxel gradients against central finite differences (h = 1e-3) of the
   *expected* alpha image, which Beer-Lambert gives in closed form: the
   expected alpha of a camera ray is 1 - exp(-sigma_t * tau) with tau the
   optical depth of the trilinear density over the leaves the tracker covers
   (FP64 quadrature), so the difference quotient of the expectation has no
   Monte Carlo noise; only the adjoint's estimate has, and it is averaged
   over enough samples to resolve 2% on average (each voxel is also held
   to its own standard error). 64 seeded random voxels among those with a
   non-negligible gradient, on the fitting scene of the C3 workload.
2. A few Adam steps of the fit lower the image loss.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _segments(oracle, sc, viewi, proji, w, h):
    """Leaf segments (ray, node, a, b) the tracker covers, grouped by node."""
    ys, xs = np.mgrid[0:h, 0:w]
    ro, rd, gm = oracle.camera_rays(viewi, proji, w, h, xs.ravel(), ys.ravel())
    nodes, t0, t1, cnt = oracle.traverse(sc.ob, ro, rd)
    last = sc.node_count - 1
    q = sc.ob["qval"][:, 1]
    j = np.arange(nodes.shape[1])[None, :]
    ok = (j < cnt[:, None]) & (gm[:, None] != 0)
    n = nodes.astype(np.int64)
    a = np.maximum(t0.astype(np.float64), 0.0)
    b = t1.astype(np.float64)
    ok &= (n != last) & (q[n] > 0) & (b > a)
    ray = np.broadcast_to(np.arange(len(ro))[:, None], n.shape)
    by_node = {}
    for node in np.unique(n[ok]):
        m = ok & (n == node)
        by_node[int(node)] = (ray[m], a[m], b[m])
    return ro.astype(np.float64), rd.astype(np.float64), by_node


def _corner_weights(p, ce, l):
    """Trilinear corner indices and weights of the lookup (FP64), (..., 8)."""
    ni = 0.5 * (p - (ce[:3] - ce[3])) * (l - 1) / ce[3]
    a = np.clip(np.floor(ni), 0, l - 1).astype(np.int64)
    c = np.clip(np.ceil(ni), 0, l - 1).astype(np.int64)
    f = ni - np.maximum(np.floor(ni), 0)
    idx = np.zeros(p.shape[:-1] + (8,), np.int64)
    wt = np.zeros(p.shape[:-1] + (8,))
    for v in range(8):
        dx, dy, dz = v & 1, (v >> 1) & 1, (v >> 2) & 1
        x = np.where(dx, c[..., 0], a[..., 0])
        y = np.where(dy, c[..., 1], a[..., 1])
        z = np.where(dz, c[..., 2], a[..., 2])
        idx[..., v] = (z * l + y) * l + x
        wt[..., v] = ((f[..., 0] if dx else 1 - f[..., 0]) * (f[..., 1] if dy else 1 - f[..., 1]) *
                      (f[..., 2] if dz else 1 - f[..., 2]))
    return idx, wt


def _leaf_matrix(sc, ro, rd, node, seg, steps):
    """d tau / d (brick value) for every segment of one leaf: (nseg, lvoxc^3),
    midpoint quadrature in FP64."""
    rays, a, b = seg
    u = (np.arange(steps) + 0.5) / steps
    t = a[:, None] + u[None, :] * (b - a)[:, None]
    p = ro[rays][:, None, :] + t[..., None] * rd[rays][:, None, :]
    idx, wt = _corner_weights(p, sc.ob["ce"][node].astype(np.float64), sc.lvoxc)
    wt = wt * ((b - a) / steps)[:, None, None]
    nseg = len(rays)
    key = (np.arange(nseg)[:, None, None] * sc.lvoxc3 + idx).ravel()
    return np.bincount(key, weights=wt.ravel(), minlength=nseg * sc.lvoxc3).reshape(nseg, sc.lvoxc3)


def _view_terms(oracle, sc, viewi, proji, w, h, sigma, voxels=None, h_fd=1e-3, steps=64):
    """Expected-alpha derivative of one view for every brick value, and the
    central difference quotients (step h_fd) for the values in `voxels`."""
    ro, rd, by_node = _segments(oracle, sc, viewi, proji, w, h)
    bricks = sc.pvol[1].reshape(-1, sc.lvoxc3).astype(np.float64)
    tau = np.zeros(len(ro))
    for node, seg in by_node.items():
        m = _leaf_matrix(sc, ro, rd, node, seg, steps)
        np.add.at(tau, seg[0], m @ bricks[int(sc.ob["volx"][node, 1])])
    grad = np.zeros(sc.pvol[1].size)
    fd = np.zeros(0 if voxels is None else len(voxels))
    for node, seg in by_node.items():
        m = _leaf_matrix(sc, ro, rd, node, seg, steps)
        vb = int(sc.ob["volx"][node, 1])
        tr = tau[seg[0]]
        grad[vb * sc.lvoxc3:(vb + 1) * sc.lvoxc3] += (sigma * np.exp(-sigma * tr)) @ m
        if voxels is not None:
            for k, g in enumerate(voxels):
                if g // sc.lvoxc3 == vb:
                    d = m[:, g % sc.lvoxc3]
                    fd[k] += float(np.sum(np.exp(-sigma * (tr - h_fd * d)) - np.exp(-sigma * (tr + h_fd * d)))) / (2 * h_fd)
    return grad, fd


def test_alpha_gradient_per_voxel_matches_central_differences(pkg, oracle, gpu):
    fw = pkg.configs.c3(views=4, res=64, spp=8, scattevs=0)
    sc = fw.init
    sig = fw.sigma_t
    grad = sum(_view_terms(oracle, sc, vi, pi, fw.w, fw.h, sig)[0] for vi, pi in fw.views)
    rng = np.random.default_rng(64)
    voxels = rng.choice(np.nonzero(grad > 0.2 * grad.max())[0], size=64, replace=False)
    fd = sum(_view_terms(oracle, sc, vi, pi, fw.w, fw.h, sig, voxels)[1] for vi, pi in fw.views)
    # GPU adjoint of the summed alpha image, averaged over seeds.
    spp = 8192
    adj_l = np.zeros((fw.h, fw.w, 4), np.float32)
    adj_l[..., 3] = 1.0
    adj_e = np.zeros_like(adj_l)
    from droplet_render_amd.fit import _kernel
    ks = [_kernel(sc, vi, pi, fw.w, fw.h, 0, sig) for vi, pi in fw.views]
    est = []
    for s in range(16):
        g = np.zeros(sc.pvol[1].size)
        for v, k in enumerate(ks):
            g += k.adjoint(0, 0, fw.w, fw.h, spp, 100 * s + v, 0, adj_l, adj_e).astype(np.float64)
        est.append(g[voxels] / spp)
    for k in ks:
        k.Destroy()
    est = np.array(est)
    mean = est.mean(axis=0)
    se = est.std(axis=0, ddof=1) / math.sqrt(len(est))
    rel = np.abs(mean / fd - 1)
    z = (mean - fd) / se
    print("per-voxel |adjoint/FD - 1|: mean %.4f median %.4f max %.4f; SE/FD max %.4f; z max %.2f mean %.2f"
          % (rel.mean(), np.median(rel), rel.max(), (se / fd).max(), np.abs(z).max(), z.mean()))
    assert np.abs(fd / grad[voxels] - 1).max() < 1e-4   # the difference quotient is the expectation's derivative
    assert rel.mean() < 0.02 and np.median(rel) < 0.02, (rel.mean(), np.median(rel))
    assert np.abs(z).max() < 5.0 and abs(z.mean()) < 4.0 / math.sqrt(64), (np.abs(z).max(), z.mean())


def test_density_fit_lowers_the_loss(pkg, gpu):
    from droplet_render_amd.fit import DensityFit
    fw = pkg.configs.c3(n=64, views=4, res=48, spp=8, scattevs=3)
    fit = DensityFit(fw, target_spp=64)
    rho0 = gpu.tensor(fw.init.pvol[1], device="cuda")
    l0 = float(fit.loss(rho0, seed=777).detach())
    rho, losses = fit.run(6, lr=0.05, seed=1)
    l1 = float(fit.loss(rho, seed=777).detach())
    fit.destroy()
    assert all(np.isfinite(losses))
    assert l1 < 0.9 * l0, (l0, l1, losses)

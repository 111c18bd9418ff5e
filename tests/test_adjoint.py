"""Pins of the oracle's adjoint (CPU): the derivative of the expected image
with respect to the FOG brick values, majorants held fixed.

* Beer-Lambert: camera-path alpha has expectation 1 - exp(-sigma_t tau(rho));
  along the scaling direction u = rho its derivative is sigma_t tau exp(-sigma_t tau),
  tau the FP64 optical depth of the trilinear density (the homogeneous-slab
  derivative -d exp(-sigma d) of the transmittance).
* Central finite differences with common random numbers (same seeds) of the
  full radiance (scattering, shadow rays, sky), under majorants raised 25%
  so the perturbed densities stay below them.
Both are Monte Carlo: the criterion is a z-score over independent seeds.
"""
import math

import numpy as np
import pytest


def zscore(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    se = math.sqrt(a.var(ddof=1) / len(a) + (b.var(ddof=1) / len(b) if len(b) > 1 else 0.0))
    return (a.mean() - b.mean()) / se


def test_alpha_gradient_matches_beer_lambert(pkg, oracle):
    from test_oracle_functions import optical_depth
    n = 32
    g = np.ones((n, n, n), np.float32)
    wl = pkg.configs.Workload("cube", g, pkg.Scene.from_grid(g), 16, 16, 128, 0, 3.0)
    adj_l = np.zeros((16, 16, 4), np.float32)
    adj_l[..., 3] = 1.0
    adj_e = np.zeros_like(adj_l)
    rho = wl.scene.pvol[1].astype(np.float64)
    vals = [float(oracle.adjoint(wl, 100 + s, adj_l, adj_e).astype(np.float64) @ rho) for s in range(8)]
    ys, xs = np.mgrid[0:16, 0:16]
    ro, rd, _ = oracle.camera_rays(wl.viewi, wl.proji, 16, 16, xs.ravel(), ys.ravel())
    nodes, t0, t1, cnt = oracle.traverse(wl.scene.ob, ro, rd)
    exact = 0.0
    for i in range(len(ro)):
        tau = optical_depth(wl.scene, ro[i], rd[i], nodes[i], t0[i], t1[i], cnt[i])
        exact += wl.spp * 3.0 * tau * math.exp(-3.0 * tau)
    z = (np.mean(vals) - exact) / (np.std(vals, ddof=1) / math.sqrt(len(vals)))
    assert abs(z) < 4.0, (np.mean(vals), exact, z)
    assert abs(np.mean(vals) / exact - 1) < 0.02


def _raised(pkg, wl):
    sc = wl.scene
    ob = sc.ob.copy()
    q = ob["qval"][:, 1]
    q[q > 0] *= np.float32(1.25)
    ob["qval"][:, 1] = q
    return ob


@pytest.mark.parametrize("direction", ["scale", "random"])
def test_radiance_gradient_matches_central_differences(pkg, oracle, direction):
    wl = pkg.configs.tiny(n=32, w=24, h=24, spp=96, scattevs=4)
    ob = _raised(pkg, wl)
    base = pkg.Scene(ob, wl.scene.pvol[1].copy(), wl.scene.lvoxc, wl.scene.mlevel)
    rng = np.random.default_rng(0)
    u = base.pvol[1].copy() if direction == "scale" else \
        (base.pvol[1] * rng.uniform(0, 1, base.pvol[1].size)).astype(np.float32)
    gl = np.zeros((24, 24, 4), np.float32)
    gl[..., :3] = 1000.0       # the sun's shadow-ray radiance is ~1e-4 per path
    gl[..., 3] = 0.01
    ge = np.zeros_like(gl)
    ge[..., :3] = 1.0
    h = 0.05

    def loss(pvol, seed):
        wl.scene = pkg.Scene(ob, pvol, base.lvoxc, base.mlevel)
        r = oracle.render(wl, seed=seed)
        return float((r[0].astype(np.float64) * gl).sum() + (r[1].astype(np.float64) * ge).sum())
    fd, ad = [], []
    for s in range(8):
        fd.append((loss(base.pvol[1] + h * u, s) - loss(base.pvol[1] - h * u, s)) / (2 * h))
        wl.scene = base
        ad.append(float(oracle.adjoint(wl, s, gl, ge).astype(np.float64) @ u.astype(np.float64)))
    z = zscore(ad, fd)
    assert abs(z) < 4.0, (np.mean(ad), np.mean(fd), z)
    assert abs(np.mean(ad) / np.mean(fd) - 1) < 0.1


def test_adjoint_is_linear_in_the_image_gradient(pkg, oracle):
    """Same keys: the adjoint of a + b equals the adjoint of a plus that of b."""
    wl = pkg.configs.tiny(n=32, w=16, h=16, spp=16, scattevs=3)
    rng = np.random.default_rng(1)
    a = [rng.normal(size=(16, 16, 4)).astype(np.float32) for _ in range(4)]
    ga = oracle.adjoint(wl, 3, a[0], a[1]).astype(np.float64)
    gb = oracle.adjoint(wl, 3, a[2], a[3]).astype(np.float64)
    gab = oracle.adjoint(wl, 3, a[0] + a[2], a[1] + a[3]).astype(np.float64)
    assert np.allclose(gab, ga + gb, rtol=1e-3, atol=1e-3 * np.abs(gab).max())
    # only leaves a ray can reach get gradients, and the never-resampled last leaf none
    last = wl.scene.ob["volx"][-1, 1]
    assert not gab.reshape(-1, 512)[last].any()

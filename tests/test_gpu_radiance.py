"""The gfx950 kernels against the FP64 radiance expectations directly (not
through the oracle): environment furnace, single scattering toward a backlit
sun with a constant environment, the Shadow pass, and MapEnv's direction
mapping (see tests/radiance_cases.py and oracle/transport64.py)."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import radiance_cases as RC  # noqa: E402

pytestmark = pytest.mark.gpu


def _render(wl, seed, x0=0, y0=0, tilex=None, tiley=None):
    tilex, tiley = tilex or wl.w, tiley or wl.h
    k = wl.kernel(tilex, tiley)
    k.set_seed(seed, 0)
    k.Render(x0, y0, tilex, tiley, wl.spp)
    out = k.phb + k.sumsq
    k.Destroy()
    return out


def test_gpu_env_furnace(pkg, gpu):
    wl = RC.furnace_workload(pkg)
    print("furnace", wl.w, wl.h, wl.spp, flush=True)
    img = _render(wl, 131)
    RC.check_furnace(img[0], img[1], wl, RC.furnace_expectation(pkg, wl))


def test_gpu_single_scattering(pkg, gpu):
    wl = RC.single_workload(pkg)
    x0, y0, tw, th = RC.SINGLE["tile"]
    print("single", x0, y0, tw, th, wl.spp, flush=True)
    img = _render(wl, 141, x0, y0, tw, th)
    RC.check_single(img, wl, RC.single_expectation(pkg, wl))


def test_gpu_shadow_pass_visibility(pkg, gpu):
    wl, depth = RC.shadow_workload(pkg)
    k = pkg.RenderKernel()
    k.Initialize(wl.scene, wl.viewi, wl.proji, wl.phase, wl.env, depth, wl.scattevs, wl.msigmas, wl.msigmaa, wl.w,
                 wl.h, wl.w, wl.h, 0, wl.lights)
    k.set_seed(113, 0)
    k.Shadow(0, 0, wl.w, wl.h, wl.spp)
    out = k.phb[0]
    k.Destroy()
    RC.check_shadow(out, wl, RC.shadow_expectation(pkg, wl, depth))


def test_gpu_map_env_matches_oracle_and_closed_form(pkg, oracle, gpu):
    rng = np.random.default_rng(6)
    v = rng.normal(size=(4000, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    h, w = 12, 20
    ii, jj = np.mgrid[0:h, 0:w].astype(np.float64)
    tex = np.stack([0.2 + 0.05 * ii, 0.1 + 0.03 * jj, 0.4 + 0.1 * np.cos(ii) * np.sin(jj), np.ones_like(ii)], -1)
    tex = tex.astype(np.float32)
    g = pkg.probe.map_env(tex, v).astype(np.float64)
    c = oracle.map_env(tex, v).astype(np.float64)
    assert np.abs(g - c).max() <= 2e-6 * np.abs(c).max()
    m = (np.abs(v[:, 2]) < 0.999) & ~((v[:, 0] > 0) & (np.abs(v[:, 1]) < 1e-3))
    row = h * np.arccos(-v[m, 2]) / np.pi
    col = w * (0.5 + np.arctan2(v[m, 1], -v[m, 0]) / (2 * np.pi))
    assert np.abs(g[m, 0] / (0.2 + 0.05 * row) ** 2.2 - 1).max() < 2e-5
    assert np.abs(g[m, 1] / (0.1 + 0.03 * col) ** 2.2 - 1).max() < 2e-5

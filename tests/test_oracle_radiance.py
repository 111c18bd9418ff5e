"""Pins of the oracle's radiance estimator (no GPU needed).

The same-key GPU tests prove that the kernels equal the oracle; these prove
that the oracle's images equal what the reference's estimator gives, without
using the oracle's tracker: FP64 quadratures over the brick density field
(oracle/transport64.py) for an environment furnace, single scattering toward
a backlit sun with a constant environment (MIS weights with the reference's
sun pdf, the albedo factor, escape light at r > 0 only, the dropped shadow-ray
sky), the Shadow pass's normalised visibility, MapEnv's direction mapping,
and the camera's pixel-to-ray map. Each expectation is also shown to reject
an image made with a plausible misreading of the reference.
"""
import math
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import radiance_cases as RC  # noqa: E402


def test_field_matches_leaf_trilinear(pkg):
    """The gathered FP64 field is fp64.trilinear of each point's leaf brick."""
    from oracle import transport64 as T
    wl = pkg.configs.tiny(n=32)
    f = T.Field(wl.scene)
    p = np.random.default_rng(3).uniform(-0.1, 1.1, (20000, 3))
    assert np.allclose(f.density(p), f.density_by_leaf(p), rtol=1e-13, atol=1e-15)
    assert (f.density(p) > 0).mean() > 0.05


def test_env_furnace(pkg, oracle):
    """sigma_a = 0, constant MapEnv c, negligible sun cone, 128 events: the
    environment image is c^2.2 times the FP64 scatter probability; camera
    rays that escape add nothing (src/kernel.cpp:325)."""
    wl = RC.furnace_workload(pkg)
    img = oracle.render(wl, seed=31)
    p = RC.furnace_expectation(pkg, wl)
    RC.check_furnace(img[0], img[1], wl, p)
    # a misreading that lights escaping camera rays fails
    bad = img[1].copy()
    miss = p < 1e-12
    bad[miss, :3] = wl.spp * np.float32(RC.FURNACE["env_c"]) ** np.float32(2.2)
    with pytest.raises(AssertionError):
        RC.check_furnace(img[0], bad, wl, p)


@pytest.fixture(scope="module")
def single(pkg, oracle):
    wl = RC.single_workload(pkg)
    x0, y0, tw, th = RC.SINGLE["tile"]
    img = oracle.render(wl, seed=41, x0=x0, y0=y0, tilex=tw, tiley=th)
    return wl, img


def test_single_scattering_matches_quadrature(pkg, single):
    """scattevs = 1, HG, albedo 0.6, backlit sun cone of 0.3 rad, constant
    env: lights and environment sums against the FP64 quadrature of the
    reference's recursion (src/kernel.cpp:387-430, KernelSampler.cpp:326-345)."""
    wl, img = single
    expect = RC.single_expectation(pkg, wl)
    assert expect[..., 0].min() > 0.01 and expect[..., 1].min() > 0.05
    RC.check_single(img, wl, expect)


@pytest.mark.parametrize("variant", [
    dict(sun_pdf=1.0 / (2.0 * math.pi * (1.0 - math.cos(np.float32(0.3))))),   # the uniform-cone pdf
    dict(albedo_power=2),                                                        # sigma_s/sigma_e twice
    dict(albedo_power=0),                                                        # albedo dropped
    dict(shadow_env=True),                                                       # sky2 kept
])
def test_single_scattering_rejects_misreadings(pkg, single, variant):
    """The oracle's image fails the expectation of each plausible misreading
    of the reference's MIS combination, so the pin above has teeth."""
    wl, img = single
    with pytest.raises(AssertionError):
        RC.check_single(img, wl, RC.single_expectation(pkg, wl, **variant))


def test_shadow_pass_visibility(pkg, oracle):
    """RenderKernel::Shadow (src/kernel.cpp:569-607): per sample, the lights'
    radiance of an escaping sun-sampled ray from the surface point, divided by
    lights[0]'s colour; its mean is the cone-averaged FP64 transmittance
    (secondary rays restart at their leaf's entry). Untraced pixels are 1."""
    wl, depth = RC.shadow_workload(pkg)
    out = oracle.shadow(wl, 13, depth)
    vis = RC.shadow_expectation(pkg, wl, depth)
    RC.check_shadow(out, wl, vis)
    # transmittance from the surface point instead of the leaf entry is another image
    from oracle import transport64 as T
    f = T.Field(wl.scene)
    S = RC.SHADOW
    ys, xs = np.mgrid[0:wl.h, 0:wl.w]
    d = T.camera_dirs_pinhole(RC.EYE, RC.TARGET, RC.FOV, wl.w, wl.h, xs.ravel(), ys.ravel())
    dep = depth.reshape(-1).astype(np.float64)
    m = dep < 1e29
    pts = RC.EYE[None, :] + d[m] * dep[m, None]
    sun = np.repeat(wl.lights[0].direction[None, :].astype(np.float64), len(pts), axis=0)
    naive = np.ones(len(d))
    naive[m] = np.exp(-S["sigma_t"] * f.tau(pts, sun, 0.0, f.exit_t(pts, sun), 256))
    assert np.abs(naive - vis.reshape(-1)).max() > 0.05
    with pytest.raises(AssertionError):
        RC.check_shadow(out, wl, naive.reshape(wl.h, wl.w))


def test_map_env_direction_mapping(oracle):
    """MapEnv (src/KernelSampler.cpp:384-419): rows run from the nadir (-z,
    row 0) to the zenith over the polar angle, columns over the azimuth
    measured from -x toward +y with -x in the middle column; bilinear
    interpolation reproduces textures linear in the row or column index
    exactly (also where the last cell extrapolates), then pow 2.2."""
    rng = np.random.default_rng(5)
    v = rng.normal(size=(4000, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    v = v[(np.abs(v[:, 2]) < 0.999) & ~((v[:, 0] > 0) & (np.abs(v[:, 1]) < 1e-3))]
    h, w = 12, 20
    polar = np.arccos(-v[:, 2])                  # 0 at the nadir
    azim = np.arctan2(v[:, 1], -v[:, 0])         # 0 toward -x, +pi/2 toward +y
    row = h * polar / np.pi
    col = w * (0.5 + azim / (2 * np.pi))
    ii, jj = np.mgrid[0:h, 0:w].astype(np.float64)
    tex = np.stack([0.2 + 0.05 * ii, 0.1 + 0.03 * jj, np.full_like(ii, 0.55), np.ones_like(ii)], -1)
    got = oracle.map_env(tex.astype(np.float32), v).astype(np.float64)
    want = np.stack([(0.2 + 0.05 * row) ** 2.2, (0.1 + 0.03 * col) ** 2.2, np.full(len(v), 0.55 ** 2.2)], -1)
    rel = np.abs(got - want) / want
    assert rel.max() < 2e-5, rel.max()
    # named directions
    dirs = np.array([[0, 0, -1.0], [0, 0, 1.0], [-1.0, 0, 0], [0, 1.0, 0], [0, -1.0, 0]])
    g = oracle.map_env(tex.astype(np.float32), dirs)[:, :2].astype(np.float64) ** (1 / 2.2)
    assert np.allclose(g[0, 0], 0.2, rtol=1e-5) and np.allclose(g[1, 0], 0.2 + 0.05 * h, rtol=1e-5)
    assert np.allclose(g[2, 1], 0.1 + 0.03 * w / 2, rtol=1e-5)
    assert np.allclose(g[3, 1], 0.1 + 0.03 * w * 0.75, rtol=1e-5)
    assert np.allclose(g[4, 1], 0.1 + 0.03 * w * 0.25, rtol=1e-5)


@pytest.mark.parametrize("w,h", [(7, 5), (64, 48), (33, 21)])
def test_camera_rays_match_pinhole(pkg, oracle, w, h):
    """Primary rays from the inverted view/projection matrices equal a
    pinhole camera built from its eye, target and field of view, with the
    reference's flipped NDC signs (kernel.cpp:462-463): pixel (0, 0) looks
    toward the camera's right and up; odd images have a centre pixel that
    looks straight at the target."""
    from oracle import transport64 as T
    vi, pi = pkg.camera.camera(RC.EYE, RC.TARGET, RC.FOV, w, h)
    ys, xs = np.mgrid[0:h, 0:w]
    ro, rd, gm = oracle.camera_rays(vi, pi, w, h, xs.ravel(), ys.ravel())
    d = T.camera_dirs_pinhole(RC.EYE, RC.TARGET, RC.FOV, w, h, xs.ravel(), ys.ravel())
    assert np.abs(rd - d).max() < 1e-6
    assert np.allclose(ro, RC.EYE[None, :], atol=1e-7) and gm.all()
    f = (RC.TARGET - RC.EYE) / np.linalg.norm(RC.TARGET - RC.EYE)
    right = np.cross(f, (0.0, 0.0, 1.0))
    assert np.dot(rd[0], right) > 0 and rd[0][2] > f[2]
    if w % 2 and h % 2:
        c = (h // 2) * w + w // 2
        assert np.abs(rd[c] - f).max() < 1e-6

"""Radiance pins shared by the oracle tests and the GPU tests.

Each case is a small scene whose expected image the FP64 quadratures of
oracle/transport64.py give without the oracle's tracker: an environment
furnace, single scattering toward a backlit sun with a constant environment,
and the Shadow pass. The checks are z-tests of per-pixel Monte Carlo means
against those expectations, with an allowance for the quadrature's own error.
"""
import math

import numpy as np

EYE = np.array((0.5, -1.2, 0.6))
TARGET = np.array((0.5, 0.5, 0.5))
FOV = 70.0   # wide enough that the image corners miss the cube

_CACHE = {}


def cube_scene(pkg, n=32):
    """Unit cube of density 1 (the bricks fall off to 1/2 at its faces, where
    the resampling meets the empty background); the last leaf is transparent."""
    key = ("cube", n)
    if key not in _CACHE:
        g = np.ones((n, n, n), np.float32)
        _CACHE[key] = (g, pkg.Scene.from_grid(g))
    return _CACHE[key]


def view(pkg, w, h):
    vi, pi = pkg.camera.camera(EYE, TARGET, FOV, w, h)
    return dict(viewi=vi, proji=pi)


def const_env(pkg, c):
    tex = np.full((8, 16, 4), c, np.float32)
    tex[..., 3] = 1.0
    return pkg.MapEnv(tex)


# ------------------------------------------------------------------ furnace --
FURNACE = dict(w=24, h=24, spp=2048, sigma_t=1.0, env_c=0.7, scattevs=128)


def furnace_workload(pkg):
    """No absorption, constant environment c, sun cone of 1e-3 rad: every
    camera path that scatters eventually escapes along a phase-sampled
    continuation whose MIS weight is 1 (the sun cone is too small to be hit),
    so the environment image is c^2.2 times the scatter probability."""
    F = FURNACE
    g, sc = cube_scene(pkg)
    sun = pkg.SunLight(pkg.camera.sun_direction(30.0, 0.0), angle=1e-3)
    return pkg.configs.Workload("furnace", g, sc, F["w"], F["h"], F["spp"], F["scattevs"], F["sigma_t"], albedo=1.0,
                                env=const_env(pkg, F["env_c"]), lights=[sun], **view(pkg, F["w"], F["h"]))


def furnace_expectation(pkg, wl):
    """Per pixel: expected scatter probability 1 - exp(-sigma_t tau) (FP64)."""
    from oracle import transport64 as T
    key = ("furnace", wl.w, wl.h)
    if key not in _CACHE:
        field = T.Field(wl.scene)
        ys, xs = np.mgrid[0:wl.h, 0:wl.w]
        d = T.camera_dirs_pinhole(EYE, TARGET, FOV, wl.w, wl.h, xs.ravel(), ys.ravel())
        a = np.array([field.camera_collisions(EYE, d[i], wl.sigma_t, 512)[1].sum() for i in range(len(d))])
        _CACHE[key] = a.reshape(wl.h, wl.w)
    return _CACHE[key]


def check_furnace(img_l, img_e, wl, p):
    """img_*: (h, w, 4) sums. Environment sums against c^2.2 * p per pixel
    (binomial error), exact zeros where no camera ray meets the cube."""
    n = wl.spp
    c22 = np.float64(FURNACE["env_c"]) ** 2.2
    alpha = img_e[..., 3].astype(np.float64)
    env = img_e[..., :3].astype(np.float64)
    assert np.array_equal(img_l[..., 3], img_e[..., 3])
    miss = p < 1e-12
    assert miss.any() and (~miss).sum() > 100
    assert not env[miss].any() and not alpha[miss].any()
    # per path: c^2.2 or 0, never more
    assert np.all(env <= alpha[..., None] * c22 * (1 + 1e-5) + 1e-6)
    hit = ~miss
    for ch in range(3):
        m = env[..., ch][hit] / n
        mu = c22 * p[hit]
        se = c22 * np.sqrt(np.maximum(p[hit] * (1 - p[hit]), 1e-12) / n)
        z = (m - mu) / se
        assert np.abs(z).max() < 5.0, (ch, np.abs(z).max())
        assert abs(z.mean()) < 4.0 / math.sqrt(len(z)), (ch, z.mean())
    za = (alpha[hit] / n - p[hit]) / np.sqrt(np.maximum(p[hit] * (1 - p[hit]), 1e-12) / n)
    assert np.abs(za).max() < 5.0 and abs(za.mean()) < 4.0 / math.sqrt(len(za))


# -------------------------------------------------------- single scattering --
SINGLE = dict(w=16, h=16, tile=(7, 7, 2, 2), spp=262144, sigma_t=1.5, albedo=0.6, g=0.5, angle=0.3,
              color=(1.0, 0.5, 0.25), env_c=0.8)
SINGLE_QUAD_REL = 4e-3   # quadrature allowance (measured against a 2x finer rule: < 2e-3)


def single_workload(pkg):
    """One scattering event (scattevs = 1), HG phase, albedo 0.6, a sun of
    half-angle 0.3 rad straight ahead of the camera (backlit: the MIS
    weights are large) and a constant environment."""
    S = SINGLE
    g, sc = cube_scene(pkg)
    d = TARGET - EYE
    d = d / np.linalg.norm(d)
    sun = pkg.SunLight(d, color=S["color"], angle=S["angle"])
    return pkg.configs.Workload("single", g, sc, S["w"], S["h"], S["spp"], 1, S["sigma_t"], albedo=S["albedo"],
                                phase=pkg.HGPhase(S["g"]), env=const_env(pkg, S["env_c"]), lights=[sun],
                                **view(pkg, S["w"], S["h"]))


def single_expectation(pkg, wl, **variant):
    """Per pixel of the tile: (E[L_sun]/colour, E[L_env]/c^2.2, E[alpha]);
    `variant` (SingleScatter.values) restates a plausible misreading."""
    from oracle import transport64 as T
    key = ("single",)
    if key not in _CACHE:
        S = SINGLE
        field = T.Field(wl.scene)
        x0, y0, tw, th = S["tile"]
        ys, xs = np.mgrid[y0:y0 + th, x0:x0 + tw]
        d = T.camera_dirs_pinhole(EYE, TARGET, FOV, wl.w, wl.h, xs.ravel(), ys.ravel())
        sun = wl.lights[0].direction.astype(np.float64)
        ca = math.cos(np.float32(S["angle"]))
        sig_s, sig_a = wl.msigmas, wl.msigmaa
        _CACHE[key] = [T.SingleScatter(field, EYE, d[i], sig_s, sig_a, S["g"], sun, ca, cells=64, n_mu=6, n_phi=16,
                                       sphere=(24, 24), steps=96) for i in range(len(d))]
    x0, y0, tw, th = SINGLE["tile"]
    return np.array([q.values(**variant) for q in _CACHE[key]]).reshape(th, tw, 3)


def check_single(img, wl, expect):
    """img: [lights, env, sumsq lights, sumsq env] (th, tw, 4) of the tile."""
    S = SINGLE
    n = wl.spp
    c22 = np.float64(S["env_c"]) ** 2.2
    col = np.array(S["color"], np.float64)
    zs = []
    for b, scale, e in [(0, col, expect[..., 0]), (1, np.full(3, c22), expect[..., 1])]:
        for ch in range(3):
            m = img[b][..., ch].astype(np.float64) / n
            var = np.maximum(img[2 + b][..., ch].astype(np.float64) / n - m * m, 0.0) / n
            mu = scale[ch] * e
            se = np.sqrt(var + (SINGLE_QUAD_REL * mu) ** 2)
            z = (m - mu) / se
            zs.append(z)
            assert np.abs(z).max() < 4.5, (b, ch, m, mu, z)
            # aggregate over the tile
            zt = (m.sum() - mu.sum()) / math.sqrt(var.sum() + (SINGLE_QUAD_REL * mu.sum()) ** 2)
            assert abs(zt) < 4.0, (b, ch, m.sum(), mu.sum(), zt)
    a = img[0][..., 3].astype(np.float64) / n
    ea = expect[..., 2]
    za = (a - ea) / np.sqrt(np.maximum(ea * (1 - ea), 1e-12) / n + (1e-3 * ea) ** 2)
    assert np.abs(za).max() < 4.5, za
    return zs


# ------------------------------------------------------------------- Shadow --
SHADOW = dict(w=16, h=16, spp=4096, sigma_t=2.0, angle=0.05, color=(2.0, 1.0, 0.5), depth=1.75)


def shadow_workload(pkg):
    S = SHADOW
    g, sc = cube_scene(pkg)
    sun = pkg.SunLight(pkg.camera.sun_direction(30.0, 40.0), color=S["color"], angle=S["angle"])
    wl = pkg.configs.Workload("shadow", g, sc, S["w"], S["h"], S["spp"], 4, S["sigma_t"], albedo=0.99,
                              lights=[sun], **view(pkg, S["w"], S["h"]))
    depth = np.full((S["h"], S["w"]), S["depth"], np.float32)
    depth[:, :3] = 1e30        # no surface: untraced, fully visible
    return wl, depth


def shadow_expectation(pkg, wl, depth):
    """Per pixel: expected sun visibility of the surface point at the depth:
    the mean over the sun cone of the secondary-ray escape probability."""
    from oracle import transport64 as T
    key = ("shadow",)
    if key not in _CACHE:
        S = SHADOW
        field = T.Field(wl.scene)
        ys, xs = np.mgrid[0:wl.h, 0:wl.w]
        d = T.camera_dirs_pinhole(EYE, TARGET, FOV, wl.w, wl.h, xs.ravel(), ys.ravel())
        dep = depth.reshape(-1).astype(np.float64)
        sun = wl.lights[0].direction.astype(np.float64)
        ca = math.cos(np.float32(S["angle"]))
        cd, cw = T.cone_quadrature(sun, ca, 4, 12)
        vis = np.ones(len(d))
        m = dep < 1e29
        pts = EYE[None, :] + d[m] * dep[m, None]
        k = len(pts)
        Tm = field.transmittance_secondary(np.repeat(pts, len(cd), axis=0), np.tile(cd, (k, 1)), S["sigma_t"],
                                           256).reshape(k, len(cd))
        vis[m] = Tm @ cw / cw.sum()
        _CACHE[key] = vis.reshape(wl.h, wl.w)
    return _CACHE[key]


def check_shadow(out, wl, vis):
    """out: (h, w, 4) Shadow sums: visibility per channel (divided by the sun
    colour), alpha = sample count."""
    n = wl.spp
    assert np.all(out[..., 3] == n)
    v = out[..., :3].astype(np.float64) / n
    assert np.array_equal(v[..., 0], v[..., 1]) and np.array_equal(v[..., 0], v[..., 2])
    se = np.sqrt(np.maximum(vis * (1 - vis), 1e-12) / n + (2e-4 * vis) ** 2)
    z = (v[..., 0] - vis) / se
    assert np.abs(z).max() < 5.0, np.abs(z).max()
    assert abs(z.mean()) < 4.0 / math.sqrt(z.size), z.mean()
    assert vis.min() < 0.5 and (vis > 0.999).any()

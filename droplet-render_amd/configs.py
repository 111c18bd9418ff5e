"""Synthetic workloads of the benchmark configurations.

Each builder returns a `Workload`: the density grid, the octree scene, camera,
samplers and render settings. Densities are procedural (no datasets exist
offline); the lighting follows the reference's defaults (sun colour (1,1,1),
intensity 1, angle 0.010 rad, addon/panel.py:360-362) with the sun at 30
degrees elevation, Mie phase, and a Hosek sky of turbidity 3 and ground albedo 0.1.
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .camera import camera, sun_direction
from .sampler import HosekSky, MiePhase, SunLight
from .scene import Scene

EYE = (0.5, -1.2, 0.6)
TARGET = (0.5, 0.5, 0.5)
FOV_Y = 40.0


def synth_fbm(n, seed, octaves=4, freq=4.0, lacunarity=2.0, gain=0.5, threshold=0.3):
    """fBm gradient-noise cloud: max(f - threshold, 0) normalised to max 1."""
    out = np.empty((n, n, n), dtype=np.float32)
    prm = np.array([octaves, freq, lacunarity, gain, threshold], dtype=np.float32)
    L.check(L.lib().drk_synth_cloud(0, n, int(seed), L.fptr(prm), len(prm), L.fptr(out)))
    return out


def synth_spheres(n, seed, spheres=64, cmin=0.2, cmax=0.8, rmin=0.05, rmax=0.2, amp=0.15, octaves=5, freq=4.0,
                  band=None):
    """Union of seeded spheres displaced by fBm; density ramps over `band` (default 2 voxels)."""
    out = np.empty((n, n, n), dtype=np.float32)
    band = 2.0 / n if band is None else band
    prm = np.array([spheres, cmin, cmax, rmin, rmax, amp, octaves, freq, band], dtype=np.float32)
    L.check(L.lib().drk_synth_cloud(1, n, int(seed), L.fptr(prm), len(prm), L.fptr(out)))
    return out


@dataclass
class Workload:
    name: str
    grid: np.ndarray
    scene: Scene
    w: int
    h: int
    spp: int
    scattevs: int
    sigma_t: float
    albedo: float = 0.99
    phase: object = field(default_factory=MiePhase)
    env: object = field(default_factory=lambda: HosekSky(3.0, 0.1))
    lights: list = field(default_factory=lambda: [SunLight(sun_direction(30.0, 0.0))])
    viewi: np.ndarray = None
    proji: np.ndarray = None

    def __post_init__(self):
        if self.viewi is None:
            self.viewi, self.proji = camera(EYE, TARGET, FOV_Y, self.w, self.h)

    @property
    def msigmas(self):
        return self.sigma_t * self.albedo

    @property
    def msigmaa(self):
        return self.sigma_t * (1.0 - self.albedo)

    @property
    def samples(self):
        return self.w * self.h * self.spp

    def kernel(self, tilex=None, tiley=None):
        from .kernel import RenderKernel
        k = RenderKernel()
        k.Initialize(self.scene, self.viewi, self.proji, self.phase, self.env, None, self.scattevs, self.msigmas,
                     self.msigmaa, tilex or self.w, tiley or self.h, self.w, self.h, 0, self.lights)
        return k


def c0(w=256, h=256, spp=16):
    """64^3 fBm cloud (4 octaves, f 4, lacunarity 2, gain 0.5, seed 1, threshold 0.3), sigma_t 8, 8 bounces."""
    g = synth_fbm(64, 1)
    return Workload("c0_fbm64", g, Scene.from_grid(g), w, h, spp, 8, 8.0)


def c1(w=1024, h=576, spp=64):
    """256^3 sphere-union cloud (64 spheres, 5-octave fBm displacement 0.15), sigma_t 16, 32 bounces."""
    g = synth_spheres(256, 1)
    return Workload("c1_spheres256", g, Scene.from_grid(g), w, h, spp, 32, 16.0)


def tiny(n=32, w=32, h=32, spp=8, scattevs=4, sigma_t=8.0, seed=3):
    """Small fBm scene for parity tests that the CPU oracle finishes in seconds."""
    g = synth_fbm(n, seed, octaves=3, freq=3.0, threshold=0.2)
    return Workload("tiny_fbm%d" % n, g, Scene.from_grid(g), w, h, spp, scattevs, sigma_t)


def _cloud_with_occupancy(n, seed, occupancy, kind):
    """An fBm (kind 0) or sphere-union (kind 1) cloud whose fraction of
    non-empty voxels is `occupancy`: the field is thresholded at the matching
    quantile and renormalised to max 1."""
    if kind == 0:
        f = synth_fbm(n, seed, threshold=-4.0)            # monotone in the raw fBm value
    else:
        f = synth_spheres(n, seed, spheres=24, band=0.5)   # ramps wide enough that few voxels saturate
    q = float(np.quantile(f, 1.0 - occupancy))
    if q >= float(f.max()):
        q = float(np.nextafter(np.float32(q), np.float32(-1.0)))
    g = np.maximum(f - q, 0.0)
    m = float(g.max())
    return (g / m if m > 0 else g).astype(np.float32)


def c2(nscenes=32, seed=2, max_n=512):
    """Batched forward workload: `nscenes` independent clouds, grid n
    log-uniform in [32, max_n] (multiples of 8), voxel occupancy uniform in
    [2%, 30%], alternating fBm / sphere union, images from {128^2, 256^2,
    512^2, 1024^2}, 16-64 spp, 16 scattering events, sigma_t 16."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(nscenes):
        n = int(round(np.exp(rng.uniform(np.log(32), np.log(max_n))) / 8.0)) * 8
        n = max(32, min(max_n, n))
        occ = float(rng.uniform(0.02, 0.30))
        res = int(rng.choice([128, 256, 512, 1024]))
        spp = int(rng.integers(16, 65))
        g = _cloud_with_occupancy(n, 1000 + i, occ, i % 2)
        out.append(Workload("c2_%02d_n%d" % (i, n), g, Scene.from_grid(g), res, res, spp, 16, 16.0))
    return out


@dataclass
class FitWorkload:
    """Differentiable density fitting: a ground-truth cloud, the scene being
    fitted (its own octree and bricks) and the ring of views."""
    gt: Workload
    init: Scene
    views: list
    w: int
    h: int
    spp: int
    scattevs: int
    sigma_t: float


def ring_views(n, elevation_deg=20.0, radius=2.6, center=(0.5, 0.5, 0.5), w=256, h=256, fov=40.0):
    """`n` cameras on a ring at `elevation_deg` around `center`, looking at it."""
    out = []
    el = np.radians(elevation_deg)
    for i in range(n):
        az = 2.0 * np.pi * i / n
        eye = (center[0] + radius * np.cos(az) * np.cos(el), center[1] + radius * np.sin(az) * np.cos(el),
               center[2] + radius * np.sin(el))
        out.append(camera(eye, center, fov, w, h))
    return out


def c3(n=128, views=16, res=256, spp=32, scattevs=8, sigma_t=8.0, seed=7, init_density=0.5):
    """Density fitting: ground truth a 128^3 fBm cloud (seed 7); the fitted
    scene starts as a uniform density 0.5 (of sigma_t) over the whole grid,
    with majorants 1 so that densities may grow to 1; 16 views on a ring at
    20 degrees elevation, 256^2, 32 spp, 8 scattering events."""
    g = synth_fbm(n, seed)
    gt = Workload("c3_gt_fbm%d" % n, g, Scene.from_grid(g), res, res, spp, scattevs, sigma_t)
    init = Scene.from_grid(np.full((n, n, n), init_density, np.float32))
    init.raise_majorants(1.0, floor=1.0)
    return FitWorkload(gt, init, ring_views(views, w=res, h=res), res, res, spp, scattevs, sigma_t)


def c4(n=512, views=8, w=1920, h=1080, spp=16, scattevs=64, sigma_t=32.0, seed=3):
    """Large forward+adjoint: a 512^3 sparse cloud (3 seeded spheres displaced
    by 6-octave fBm; about 2% of the leaf lattice holds density), sigma_t 32,
    8 views at 1920x1080 on a 20-degree ring, 16 spp, 64 scattering events.
    Returns (workload of view 0, list of (viewi, proji) for all views)."""
    g = synth_spheres(n, seed, spheres=3, rmin=0.03, rmax=0.06, amp=0.05, octaves=6)
    vs = ring_views(views, w=w, h=h, radius=1.6, fov=30.0)
    wl = Workload("c4_spheres%d" % n, g, Scene.from_grid(g), w, h, spp, scattevs, sigma_t,
                  viewi=vs[0][0], proji=vs[0][1])
    return wl, vs

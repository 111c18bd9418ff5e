"""KernelSampler: phase functions, lights and environments.

Mirrors the objects droplet-render's render kernel is handed
(src/KernelSampler.h:6-97): `MiePhase` / `HGPhase` (PhaseFunction), `SunLight`
(BaseLight; `lights[0]` is the MIS light), `NullEnv` / `MapEnv` (BaseEnv) and
the Hosek-Wilkie sky (compiled in with USE_ARHOSEK_SKYMODEL, src/kernel.cpp:329-360),
whose turbidity and ground albedo the reference hard-codes to 7.0 / 0.15
(src/kernel.cpp:525) and which are parameters here.
"""
import ctypes as C

import numpy as np

from . import _lib as L


class MiePhase:
    kind = L.DRK_PHASE_MIE
    g1 = 0.0


class HGPhase:
    kind = L.DRK_PHASE_HG

    def __init__(self, g=0.75):
        if g == 0.0 or not (-1.0 < g < 1.0):
            raise ValueError("HGPhase: g must be in (-1,1) and non-zero")
        self.g1 = float(g)


class SunLight:
    """SunLight(direction, color*intensity, angle), src/KernelSampler.cpp:305-345."""

    def __init__(self, direction, color=(1.0, 1.0, 1.0), intensity=1.0, angle=0.010):
        d = np.asarray(direction, dtype=np.float64)
        self.direction = (d / np.linalg.norm(d)).astype(np.float32)
        self.color = (np.asarray(color, dtype=np.float32) * np.float32(intensity)).astype(np.float32)
        self.angle = float(angle)

    def to_c(self):
        s = L.Sun()
        s.direction[:] = [float(x) for x in self.direction]
        s.color[:] = [float(x) for x in self.color]
        s.angle = self.angle
        return s


class HosekSky:
    kind = L.DRK_ENV_HOSEK

    def __init__(self, turbidity=7.0, albedo=0.15):
        self.turbidity = float(turbidity)
        self.albedo = float(albedo)


class NullEnv:
    kind = L.DRK_ENV_NULL


class MapEnv:
    """Equirectangular RGBA environment (h rows of w texels)."""
    kind = L.DRK_ENV_MAP

    def __init__(self, texels):
        t = np.ascontiguousarray(texels, dtype=np.float32)
        if t.ndim != 3 or t.shape[2] != 4 or t.shape[0] < 2 or t.shape[1] < 2:
            raise ValueError("MapEnv: texels must be (h, w, 4) with h, w >= 2")
        self.texels = t
        self.h, self.w = t.shape[:2]


def sampler_desc(phase, env, lights):
    """Builds the C descriptor; the returned tuple keeps the arrays alive."""
    lights = list(lights)
    if not lights:
        raise ValueError("at least one light is required (lights[0] is sampled)")
    arr = (L.Sun * len(lights))(*[l.to_c() for l in lights])
    d = L.SamplerDesc()
    d.phase = phase.kind
    d.hg_g = float(getattr(phase, "g1", 0.0))
    d.env = env.kind
    d.sky_turbidity = float(getattr(env, "turbidity", 0.0))
    d.sky_albedo = float(getattr(env, "albedo", 0.0))
    keep = [arr]
    if env.kind == L.DRK_ENV_MAP:
        d.env_map = L.fptr(env.texels)
        d.env_w, d.env_h = env.w, env.h
        keep.append(env.texels)
    d.lights = C.cast(arr, C.POINTER(L.Sun))
    d.light_count = len(lights)
    return d, keep

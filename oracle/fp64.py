"""Float64 restatements of the path's point functions (TEST INFRASTRUCTURE ONLY).

The FP32 oracle and the gfx950 kernels are checked against these to the
tolerances the parity tests state. Each follows the reference function named:
  mie_eval       MiePhase::EvaluateRGB      src/KernelSampler.cpp:233-256
  mie_theta      MiePhase::Sample's scan    src/KernelSampler.cpp:258-273
  hg_eval        HGPhase::Evaluate          src/KernelSampler.cpp:173-177
  sky            src/kernel.cpp:329-360 with the Hosek state of src/ArHosekSkyModel.cpp:147-339
  trilinear      SampleVoxelSpace           src/kernel.cpp:96-126
"""
import numpy as np

from .pyoracle import MIE_CDF, MIE_PDF, hosek_state, sun_elevation

PDF = MIE_PDF.astype(np.float64)
CDF = MIE_CDF.astype(np.float64)


def mie_eval(ct):
    ct = np.clip(np.asarray(ct, np.float64), -1.0, 1.0)
    b = 256.0 * np.arccos(-ct) / np.pi
    c = np.clip(np.floor(b), 0, 254).astype(np.int64)
    t = (b - c)[:, None]
    return t * (PDF[c + 1] - PDF[c]) + PDF[c]


def mie_theta_fp32(u1):
    """The reference's scan evaluated in float32 for many u at once (same
    arithmetic as the C oracle, vectorised): returns theta/pi."""
    u = np.asarray(u1, np.float32)
    ub = np.float32(256.0) * u
    cdf = MIE_CDF[:, 0].astype(np.float32)
    fl = np.minimum(np.floor(ub), np.float32(255.0))
    t_lo = (ub - fl).astype(np.float32)                      # t for i <= ub
    i = np.arange(255, dtype=np.float32)
    below = i[None, :] <= ub[:, None]                        # max(ub, i) == ub
    t = np.where(below, t_lo[:, None], np.float32(0.0)).astype(np.float32)
    keys = (t * (cdf[1:256][None, :] - cdf[0:255][None, :]) + cdf[0:255][None, :]).astype(np.float32)
    m = keys < u[:, None]
    # first index where the key is not below u; the scan stops there
    first_false = np.where(m.all(axis=1), 255, np.argmin(m, axis=1))
    last = first_false - 1
    tt = t[np.arange(len(u)), np.maximum(last, 0)]
    th = np.where(last >= 0, (last.astype(np.float32) / np.float32(256.0)) + tt / np.float32(256.0), np.float32(0.0))
    return th.astype(np.float32), keys, m


def hg_eval(g, ct):
    ct = np.asarray(ct, np.float64)
    return (1.0 - g * g) / (4.0 * np.pi * (1.0 + g * g - 2.0 * g * ct) ** 1.5)


def sky_cph(rd, d):
    """The reference's cos(gamma) construction (not a dot product)."""
    rd = np.asarray(rd, np.float64)
    d = np.asarray(d, np.float64)
    sth = np.sqrt(1.0 - rd[:, 2] ** 2)
    slth = np.sqrt(1.0 - d[2] ** 2)
    rdd = rd[:, 0] / rd[:, 1]
    acr = 1.0 / np.sqrt(rdd * rdd + 1.0)
    czp = acr * (rdd * d[1] + d[0])
    return sth * slth * czp + rd[:, 2] * d[2]


def sky(direction, turbidity, albedo, rd, cfg=None, rad=None):
    if cfg is None:
        cfg, rad = hosek_state(turbidity, albedo, sun_elevation(direction))
        cfg = cfg.astype(np.float32).astype(np.float64)   # the kernel casts the state to float
        rad = rad.astype(np.float32).astype(np.float64)
    rd = np.asarray(rd, np.float64)
    cph = sky_cph(rd, direction)
    gmma = np.arccos(cph)
    gacs = np.clip(cph, 0.0, 1.0)
    thcs = np.maximum(rd[:, 2], 0.0)
    raym = gacs * gacs
    out = np.zeros((len(rd), 3))
    for i in range(3):
        c = cfg[i]
        expm = np.exp(c[4] * gmma)
        miem = (1.0 + raym) / (1.0 + c[8] * c[8] - 2.0 * c[8] * gacs) ** 1.5
        zenh = np.sqrt(thcs)
        ca = (1.0 + c[0] * np.exp(c[1] / (thcs + 0.01))) * (c[2] + c[3] * expm + c[5] * raym + c[6] * miem + c[7] * zenh)
        out[:, i] = 1e-3 * (ca * rad[i]) ** 2.2
    return out


def trilinear(brick, p, ce, lvoxc):
    b = np.asarray(brick, np.float64).reshape(lvoxc, lvoxc, lvoxc)
    p = np.asarray(p, np.float64).reshape(-1, 3)
    ce = np.asarray(ce, np.float64)
    ni = 0.5 * (p - (ce[:3] - ce[3])) * (lvoxc - 1) / ce[3]
    a = np.clip(np.floor(ni), 0, lvoxc - 1).astype(int)
    c = np.clip(np.ceil(ni), 0, lvoxc - 1).astype(int)
    f = ni - np.maximum(np.floor(ni), 0)
    out = np.zeros(len(p))
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                w = (f[:, 0] if dx else 1 - f[:, 0]) * (f[:, 1] if dy else 1 - f[:, 1]) * (f[:, 2] if dz else 1 - f[:, 2])
                x = np.where(dx, c[:, 0], a[:, 0])
                y = np.where(dy, c[:, 1], a[:, 1])
                z = np.where(dz, c[:, 2], a[:, 2])
                out += w * b[z, y, x]
    return out

"""numpy wrappers of the GPU test hooks (include/droplet_render_probe.h).

Each call runs one device function of the renderer over a batch of inputs on
the GPU; the parity tests compare the results with the CPU oracle.
"""
import ctypes as C

import numpy as np

from . import _lib as L


def _f(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a if n is None else a.reshape(n)


def _u(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def philox(ctr, key):
    ctr, key = _u(ctr).reshape(-1, 4), _u(key).reshape(-1, 2)
    out = np.zeros_like(ctr)
    L.check(L.lib().drk_probe_philox(L.uptr(ctr), L.uptr(key), len(ctr), L.uptr(out)))
    return out


def ray_stream(key0, sample, ray, key1, m):
    key0, sample, ray = _u(key0), _u(sample), _u(ray)
    out = np.zeros((len(key0), m), np.float32)
    L.check(L.lib().drk_probe_ray_stream(L.uptr(key0), L.uptr(sample), L.uptr(ray), len(key0), int(key1), int(m),
                                         L.fptr(out)))
    return out


def log_unit(n=1 << 23):
    """(logf(u), the kernels' log(u)) for u = (2i+1) 2^-24, i < n."""
    a = np.zeros(n, np.float32)
    b = np.zeros(n, np.float32)
    L.check(L.lib().drk_probe_log_unit(int(n), L.fptr(a), L.fptr(b)))
    return a, b


def mie_eval(ct):
    ct = _f(ct)
    out = np.zeros((len(ct), 3), np.float32)
    L.check(L.lib().drk_probe_mie_eval(L.fptr(ct), len(ct), L.fptr(out)))
    return out


def mie_sample(iv, u1, u2):
    iv, u1, u2 = _f(iv).reshape(-1, 3), _f(u1), _f(u2)
    out = np.zeros_like(iv)
    L.check(L.lib().drk_probe_mie_sample(L.fptr(iv), L.fptr(u1), L.fptr(u2), len(iv), L.fptr(out)))
    return out


def hg_eval(g, ct):
    ct = _f(ct)
    out = np.zeros_like(ct)
    L.check(L.lib().drk_probe_hg_eval(float(g), L.fptr(ct), len(ct), L.fptr(out)))
    return out


def hg_sample(g, iv, u1, u2):
    iv, u1, u2 = _f(iv).reshape(-1, 3), _f(u1), _f(u2)
    out = np.zeros_like(iv)
    L.check(L.lib().drk_probe_hg_sample(float(g), L.fptr(iv), L.fptr(u1), L.fptr(u2), len(iv), L.fptr(out)))
    return out


def _sun(direction, color, angle):
    s = L.Sun()
    s.direction[:] = [float(x) for x in direction]
    s.color[:] = [float(x) for x in color]
    s.angle = float(angle)
    return s


def sun(direction, color, angle, u1, u2, dirs):
    u1, u2, dirs = _f(u1), _f(u2), _f(dirs).reshape(-1, 3)
    n = len(u1)
    sd, pdf, ev = np.zeros((n, 3), np.float32), np.zeros(n, np.float32), np.zeros((n, 3), np.float32)
    s = _sun(direction, color, angle)
    L.check(L.lib().drk_probe_sun(C.byref(s), L.fptr(u1), L.fptr(u2), L.fptr(dirs), n, L.fptr(sd), L.fptr(pdf),
                                  L.fptr(ev)))
    return sd, pdf, ev


def sky(direction, turbidity, albedo, rd):
    rd = _f(rd).reshape(-1, 3)
    out = np.zeros_like(rd)
    s = _sun(direction, (1, 1, 1), 0.01)
    L.check(L.lib().drk_probe_sky(C.byref(s), float(turbidity), float(albedo), L.fptr(rd), len(rd), L.fptr(out)))
    return out


def trilinear(brick, lvoxc, ce, p, leaf_rcp=True):
    """leaf_rcp: divide by the extent through its reciprocal (the kernels'
    form when all leaves share one extent), else by the division."""
    brick, ce, p = _f(brick), _f(ce), _f(p).reshape(-1, 3)
    out = np.zeros(len(p), np.float32)
    L.check(L.lib().drk_probe_trilinear(L.fptr(brick), int(lvoxc), L.fptr(ce), L.fptr(p), len(p),
                                        int(bool(leaf_rcp)), L.fptr(out)))
    return out


def traverse(scene, ro, rd, max_leaves=256):
    ro, rd = _f(ro).reshape(-1, 3), _f(rd).reshape(-1, 3)
    n = len(ro)
    nodes = np.zeros((n, max_leaves), np.uint32)
    t0 = np.zeros((n, max_leaves), np.float32)
    t1 = np.zeros((n, max_leaves), np.float32)
    cnt = np.zeros(n, np.uint32)
    L.check(L.lib().drk_probe_traverse(scene.device(), L.fptr(ro), L.fptr(rd), n, max_leaves, L.uptr(nodes),
                                       L.fptr(t0), L.fptr(t1), L.uptr(cnt)))
    return nodes, t0, t1, cnt


def camera(viewi, proji, w, h, px, py):
    px, py = _u(px), _u(py)
    n = len(px)
    ro, rd, gm = np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32), np.zeros(n, np.uint32)
    L.check(L.lib().drk_probe_camera(L.fptr(_f(viewi, 16)), L.fptr(_f(proji, 16)), int(w), int(h), L.uptr(px),
                                     L.uptr(py), n, L.fptr(ro), L.fptr(rd), L.uptr(gm)))
    return ro, rd, gm


def map_env(texels, rd):
    """MapEnv::Evaluate then pow 2.2 on the GPU; texels (h, w, 4)."""
    tex = _f(texels)
    h, w = tex.shape[:2]
    rd = _f(rd).reshape(-1, 3)
    out = np.zeros_like(rd)
    L.check(L.lib().drk_probe_map_env(L.fptr(tex), int(w), int(h), L.fptr(rd), len(rd), L.fptr(out)))
    return out

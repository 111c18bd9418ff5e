"""Python face of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loads oracle/_build/liboracle.so (oracle.c, the C restatement of the
reference path) and adds a pure-Python restatement of the octree build
(src/scene.cpp:234-305, :676-762, :840-917). Only tests/, smoke() and
bench.py's CPU-baseline leg use this module, always as the checker.
"""
import ctypes as C
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "_build", "liboracle.so")
LIB_FAST = os.path.join(HERE, "_build", "liboracle_fast.so")   # -O3 -march=x86-64-v3 build, same numbers
REF_HOSEK = os.path.join(HERE, "_ref", "libarhosek_ref.so")

sys.path.insert(0, os.path.join(ROOT, "droplet-render_amd", "data"))
from extract_tables import load_inc  # noqa: E402  (data tables, shared as data)
sys.path.pop(0)

_F = C.POINTER(C.c_float)
_U = C.POINTER(C.c_uint32)
_D = C.POINTER(C.c_double)


class OLight(C.Structure):
    _fields_ = [("dir", C.c_float * 3), ("color", C.c_float * 3), ("cosang", C.c_float)]


class OSky(C.Structure):
    _fields_ = [("cfg", C.c_float * 27), ("rad", C.c_float * 3), ("dir", C.c_float * 3)]


class OMie(C.Structure):
    _fields_ = [("pdf", C.c_float * 768), ("cdf", C.c_float * 768)]


class OParams(C.Structure):
    _fields_ = [("ob", C.c_void_p), ("pvol_fog", _F), ("pvol_sdf", _F), ("lvoxc", C.c_uint32), ("lvoxc3", C.c_uint32),
                ("viewi", C.c_float * 16), ("proji", C.c_float * 16), ("w", C.c_uint32), ("h", C.c_uint32),
                ("scattevs", C.c_uint32), ("sigma_s", C.c_float), ("sigma_a", C.c_float), ("phase", C.c_uint32),
                ("hg_g", C.c_float), ("env", C.c_uint32), ("env_map", _F), ("env_w", C.c_uint32),
                ("env_h", C.c_uint32), ("light_count", C.c_uint32), ("lights", OLight * 8), ("sky", OSky),
                ("mie", C.POINTER(OMie)), ("depth", _F), ("key1", C.c_uint32)]


_libs = {}


def lib(fast=False):
    path = LIB_FAST if fast else LIB
    if path not in _libs:
        if not os.path.exists(path):
            raise ImportError("oracle not built: make -C oracle")
        L = C.CDLL(path)
        assert L.or_params_size() == C.sizeof(OParams), (L.or_params_size(), C.sizeof(OParams))
        assert L.or_mie_size() == C.sizeof(OMie)
        L.or_mie_theta.restype = C.c_float
        L.or_voxel.restype = C.c_float
        L.or_hg_eval.restype = C.c_float
        _libs[path] = L
    return _libs[path]


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def fp(a):
    return a.ctypes.data_as(_F)


def up(a):
    return a.ctypes.data_as(_U)


# ----------------------------------------------------------------- tables --
MIE_PDF = np.array(load_inc("mie_pdf_rgb.inc"), dtype=np.float32).reshape(256, 3)
MIE_CDF = np.array(load_inc("mie_cdf_rgb.inc"), dtype=np.float32).reshape(256, 3)
HOSEK_COEFFS = np.array(load_inc("hosek_rgb_coeffs.inc"), dtype=np.float64)
HOSEK_RAD = np.array(load_inc("hosek_rgb_radiance.inc"), dtype=np.float64)

_MIE = OMie()
_MIE.pdf[:] = [float(x) for x in MIE_PDF.reshape(-1)]
_MIE.cdf[:] = [float(x) for x in MIE_CDF.reshape(-1)]


def rng_key1(seed):
    seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    return ((seed & 0xFFFFFFFF) ^ (((seed >> 32) * 0x9E3779B9) & 0xFFFFFFFF)) & 0xFFFFFFFF


# -------------------------------------------------------------- functions --
def philox(ctr, key):
    ctr = np.ascontiguousarray(ctr, dtype=np.uint32).reshape(-1, 4)
    key = np.ascontiguousarray(key, dtype=np.uint32).reshape(-1, 2)
    out = np.zeros_like(ctr)
    lib().or_philox(up(ctr), up(key), C.c_uint32(len(ctr)), up(out))
    return out


def ray_stream(key0, sample, ray, key1, m):
    key0 = np.ascontiguousarray(key0, dtype=np.uint32)
    sample = np.ascontiguousarray(sample, dtype=np.uint32)
    ray = np.ascontiguousarray(ray, dtype=np.uint32)
    out = np.zeros((len(key0), m), dtype=np.float32)
    lib().or_ray_stream(up(key0), up(sample), up(ray), C.c_uint32(len(key0)), C.c_uint32(key1), C.c_uint32(m),
                        fp(out))
    return out


def mie_eval(ct):
    ct = f32(ct)
    out = np.zeros((len(ct), 3), np.float32)
    o = np.zeros(3, np.float32)
    for i, c in enumerate(ct):
        lib().or_mie_eval(C.byref(_MIE), C.c_float(c), fp(o))
        out[i] = o
    return out


def mie_theta(u1):
    return np.array([lib().or_mie_theta(C.byref(_MIE), C.c_float(u)) for u in f32(u1)], dtype=np.float32)


def mie_sample(iv, u1, u2):
    iv = f32(iv).reshape(-1, 3)
    out = np.zeros_like(iv)
    o = np.zeros(3, np.float32)
    for i in range(len(iv)):
        lib().or_mie_sample(C.byref(_MIE), fp(iv[i]), C.c_float(u1[i]), C.c_float(u2[i]), fp(o))
        out[i] = o
    return out


def hg_eval(g, ct):
    return np.array([lib().or_hg_eval(C.c_float(g), C.c_float(c)) for c in f32(ct)], dtype=np.float32)


def hg_sample(g, iv, u1, u2):
    iv = f32(iv).reshape(-1, 3)
    out = np.zeros_like(iv)
    o = np.zeros(3, np.float32)
    for i in range(len(iv)):
        lib().or_hg_sample(C.c_float(g), fp(iv[i]), C.c_float(u1[i]), C.c_float(u2[i]), fp(o))
        out[i] = o
    return out


def sun(direction, color, angle, u1, u2, dirs):
    u1, u2, dirs = f32(u1), f32(u2), f32(dirs).reshape(-1, 3)
    n = len(u1)
    osd, opdf, oev = np.zeros((n, 3), np.float32), np.zeros(n, np.float32), np.zeros((n, 3), np.float32)
    lib().or_sun(fp(f32(direction)), fp(f32(color)), C.c_float(angle), fp(u1), fp(u2), fp(dirs), C.c_uint32(n),
                 fp(osd), fp(opdf), fp(oev))
    return osd, opdf, oev


def hosek_state(turbidity, albedo, elevation):
    cfg, rad = np.zeros(27, np.float64), np.zeros(3, np.float64)
    lib().or_hosek_state(HOSEK_COEFFS.ctypes.data_as(_D), HOSEK_RAD.ctypes.data_as(_D), C.c_double(turbidity),
                         C.c_double(albedo), C.c_double(elevation), cfg.ctypes.data_as(_D), rad.ctypes.data_as(_D))
    return cfg.reshape(3, 9), rad


def sun_elevation(direction):
    """RenderKernel::Initialize: se = pi/2 - acosf(dir.z), in float (src/kernel.cpp:522-524)."""
    th = np.float32(math.acos(np.float32(direction[2])))
    return float(np.float32(np.float32(1.5707963267948966) - th))


def sky(direction, turbidity, albedo, rd):
    cfg, rad = hosek_state(turbidity, albedo, sun_elevation(direction))
    rd = f32(rd).reshape(-1, 3)
    out = np.zeros_like(rd)
    lib().or_sky(fp(f32(cfg.reshape(-1))), fp(f32(rad)), fp(f32(direction)), fp(rd), C.c_uint32(len(rd)), fp(out))
    return out


def voxel(brick, p, ce, lvoxc):
    brick, p, ce = f32(brick), f32(p).reshape(-1, 3), f32(ce)
    return np.array([lib().or_voxel(fp(brick), fp(p[i]), fp(ce), C.c_uint32(lvoxc)) for i in range(len(p))],
                    dtype=np.float32)


def traverse(ob, ro, rd, max_leaves=256):
    ob = np.ascontiguousarray(ob)
    ro, rd = f32(ro).reshape(-1, 3), f32(rd).reshape(-1, 3)
    n = len(ro)
    nodes = np.zeros((n, max_leaves), np.uint32)
    t0 = np.zeros((n, max_leaves), np.float32)
    t1 = np.zeros((n, max_leaves), np.float32)
    cnt = np.zeros(n, np.uint32)
    lib().or_traverse(C.c_void_p(ob.ctypes.data), fp(ro), fp(rd), C.c_uint32(n), C.c_uint32(max_leaves), up(nodes),
                      fp(t0), fp(t1), up(cnt))
    return nodes, t0, t1, cnt


def camera_rays(viewi, proji, w, h, px, py):
    px, py = np.ascontiguousarray(px, np.uint32), np.ascontiguousarray(py, np.uint32)
    n = len(px)
    ro, rd, gm = np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32), np.zeros(n, np.uint32)
    lib().or_camera(fp(f32(viewi).reshape(-1)), fp(f32(proji).reshape(-1)), C.c_uint32(w), C.c_uint32(h), up(px),
                    up(py), C.c_uint32(n), fp(ro), fp(rd), up(gm))
    return ro, rd, gm


# --------------------------------------------------------------- renderer --
def make_params(wl, seed, depth=None):
    """Fills the oracle's parameter block from a droplet_render_amd Workload
    (depth: optional w*h compositing depth)."""
    P = OParams()
    sc = wl.scene
    keep = [np.ascontiguousarray(sc.ob), f32(sc.pvol[1])]
    P.ob = keep[0].ctypes.data
    P.pvol_fog = fp(keep[1])
    if sc.pvol[0] is not None and sc.leafx[0]:
        keep.append(f32(sc.pvol[0]))
        P.pvol_sdf = fp(keep[-1])
    else:
        P.pvol_sdf = None
    P.lvoxc, P.lvoxc3 = sc.lvoxc, sc.lvoxc ** 3
    P.viewi[:] = [float(x) for x in f32(wl.viewi).reshape(-1)]
    P.proji[:] = [float(x) for x in f32(wl.proji).reshape(-1)]
    P.w, P.h = wl.w, wl.h
    P.scattevs = wl.scattevs
    P.sigma_s, P.sigma_a = float(np.float32(wl.msigmas)), float(np.float32(wl.msigmaa))
    P.phase = wl.phase.kind
    P.hg_g = float(getattr(wl.phase, "g1", 0.0))
    P.env = wl.env.kind
    if P.env == 2:
        keep.append(f32(wl.env.texels))
        P.env_map = fp(keep[-1])
        P.env_w, P.env_h = wl.env.w, wl.env.h
    P.light_count = len(wl.lights)
    for i, l in enumerate(wl.lights):
        P.lights[i].dir[:] = [float(x) for x in l.direction]
        P.lights[i].color[:] = [float(x) for x in l.color]
        P.lights[i].cosang = float(np.float32(math.cos(np.float32(l.angle))))
    if P.env == 0:
        d0 = wl.lights[0].direction
        cfg, rad = hosek_state(wl.env.turbidity, wl.env.albedo, sun_elevation(d0))
        P.sky.cfg[:] = [float(x) for x in f32(cfg.reshape(-1))]
        P.sky.rad[:] = [float(x) for x in f32(rad)]
        P.sky.dir[:] = [float(x) for x in d0]
    P.mie = C.pointer(_MIE)
    if depth is not None:
        keep.append(f32(depth).reshape(-1))
        P.depth = fp(keep[-1])
    else:
        P.depth = None
    P.key1 = rng_key1(seed)
    return P, keep


def render(wl, seed, sample_base=0, x0=0, y0=0, tilex=None, tiley=None, samples=None, threads=None, depth=None,
           fast=False, packets=False):
    """CPU render of a tile; returns [lights, env, sumsq lights, sumsq env] as (tiley, tilex, 4) float32.
    fast: the -O3 x86-64-v3 build of the same restatement.
    packets: the reference's structure (2x2 pixel packets traced in 4 lanes
    under masks, or_render_packets); the same sums bit for bit."""
    tilex = tilex or wl.w
    tiley = tiley or wl.h
    samples = samples or wl.spp
    threads = threads or os.cpu_count() or 1
    P, keep = make_params(wl, seed, depth)
    out = np.zeros((4, tiley, tilex, 4), np.float32)
    fn = lib(fast).or_render_packets if packets else lib(fast).or_render
    fn(C.byref(P), C.c_uint32(x0), C.c_uint32(y0), C.c_uint32(tilex), C.c_uint32(tiley),
       C.c_uint32(samples), C.c_uint32(sample_base), C.c_uint32(threads), fp(out))
    del keep
    return [out[0], out[1], out[2], out[3]]


def adjoint(wl, seed, adj_lights, adj_env, sample_base=0, x0=0, y0=0, tilex=None, tiley=None, samples=None):
    """CPU adjoint: dL/d(FOG brick values) given dL/d(phb sums) of both buffers."""
    tilex = tilex or wl.w
    tiley = tiley or wl.h
    samples = samples or wl.spp
    P, keep = make_params(wl, seed)
    adj = np.ascontiguousarray(np.stack([f32(adj_lights).reshape(tiley, tilex, 4),
                                         f32(adj_env).reshape(tiley, tilex, 4)]))
    n = wl.scene.leafx[1] * wl.scene.lvoxc3
    grad = np.zeros(n, np.float32)
    lib().or_adjoint(C.byref(P), C.c_uint32(x0), C.c_uint32(y0), C.c_uint32(tilex), C.c_uint32(tiley),
                     C.c_uint32(samples), C.c_uint32(sample_base), fp(adj), fp(grad), C.c_size_t(n))
    del keep
    return grad


def clip_end(proji):
    """-z/w of (0,0,1,1)*proji in float32 (src/kernel.cpp:574-575)."""
    p = f32(proji).reshape(4, 4)
    z = np.float32(np.float32(np.float32(0.0) * p[0, 2] + np.float32(0.0) * p[1, 2]) + p[2, 2]) + p[3, 2]
    w = np.float32(np.float32(np.float32(0.0) * p[0, 3] + np.float32(0.0) * p[1, 3]) + p[2, 3]) + p[3, 3]
    return float(np.float32(-z / w))


def shadow(wl, seed, depth, sample_base=0, x0=0, y0=0, tilex=None, tiley=None, samples=None):
    """CPU Shadow pass of a tile: (tiley, tilex, 4) sums (RGB visibility / sun colour, alpha = sample count)."""
    tilex = tilex or wl.w
    tiley = tiley or wl.h
    samples = samples or wl.spp
    P, keep = make_params(wl, seed, depth)
    out = np.zeros((4, tiley, tilex, 4), np.float32)
    lib().or_shadow(C.byref(P), C.c_uint32(x0), C.c_uint32(y0), C.c_uint32(tilex), C.c_uint32(tiley),
                    C.c_uint32(samples), C.c_uint32(sample_base), C.c_float(clip_end(wl.proji)), fp(out))
    del keep
    return out[0]


# ------------------------------------------------------- octree build (Python) --
def _f(x):
    return np.float32(x)


def vdb_leaves(grid, band=0.0):
    """Active 8^3 leaves of a dense (nz,ny,nx) grid in OpenVDB tree order with
    their active index bounding boxes (min, max inclusive). Active: non-zero,
    or |value| < band for a level set's narrow band."""
    nz, ny, nx = grid.shape
    out = []
    for bz in range(0, nz, 8):
        for by in range(0, ny, 8):
            for bx in range(0, nx, 8):
                blk = grid[bz:bz + 8, by:by + 8, bx:bx + 8]
                nzi = np.nonzero(np.abs(blk) < np.float32(band)) if band > 0 else np.nonzero(blk)
                if len(nzi[0]) == 0:
                    continue
                mn = (bx + nzi[2].min(), by + nzi[1].min(), bz + nzi[0].min())
                mx = (bx + nzi[2].max(), by + nzi[1].max(), bz + nzi[0].max())
                key = (bx >> 12, by >> 12, bz >> 12, (bx >> 7) & 31, (by >> 7) & 31, (bz >> 7) & 31,
                       (bx >> 3) & 15, (by >> 3) & 15, (bz >> 3) & 15)
                out.append((key, mn, mx))
    out.sort(key=lambda t: t[0])
    return [(mn, mx) for _, mn, mx in out]


def build_octree(grid, s=None, maxd=12, sdf=None):
    """Pure-Python restatement of the octree build: leaf AABBs
    (scene.cpp:676-701, buffer order SDF then FOG), root sizing (:703-716),
    BuildPath (:267-305) and the resampling with majorants and the culling of
    fog leaves inside the surface (:870-908), with float32 arithmetic where
    the reference uses float. `sdf`: a signed-distance grid standing for the
    level set (narrow band 4 voxels, values held to +-band, band outside).
    Returns (nodes list of dicts, FOG bricks (leafx,lvoxc^3), lvoxc, mlevel),
    plus the SDF bricks when `sdf` is given."""
    ref = grid if grid is not None else sdf
    nz, ny, nx = np.asarray(ref).shape
    grid = None if grid is None else np.asarray(grid, dtype=np.float32)
    sdf = None if sdf is None else np.asarray(sdf, dtype=np.float32)
    s = _f(s if s else 1.0 / nx)
    band = _f(_f(4.0) * s)

    def aabbs(g, bnd):
        out = []
        for mn, mx in vdb_leaves(g, bnd):
            c, e = [], []
            for a in range(3):
                dimi = _f(mx[a] - mn[a] + 1)
                e.append(_f(_f(_f(0.5) * dimi) * s))
                posi = (float(mn[a]) + float(mx[a])) * 0.5
                c.append(_f(posi * float(s)))
            out.append((c, e))
        return out
    boxes_b = [aabbs(sdf, float(band)) if sdf is not None else [], aabbs(grid, 0.0) if grid is not None else []]
    boxes = boxes_b[0] + boxes_b[1]
    lo = [_f(np.finfo(np.float32).max)] * 3
    hi = [_f(-np.finfo(np.float32).max)] * 3
    for c, e in boxes:
        for a in range(3):
            lo[a] = min(_f(c[a] - e[a]), lo[a])
            hi[a] = max(_f(c[a] + e[a]), hi[a])
    cr = [_f(_f(0.5) * _f(hi[a] + lo[a])) for a in range(3)]
    er = [_f(_f(0.5) * _f(hi[a] - lo[a])) for a in range(3)]
    amax = max(max(er[0], er[1]), er[2])
    lvc = _f(8.0)
    d = _f(_f(2.0) * amax)
    mlevel = int(math.ceil(np.float32(np.log(np.float32(d / _f(lvc * s)))) / _f(0.69315)))
    if mlevel > maxd:
        lvc = _f(lvc * _f(2.0 ** (mlevel - maxd)))
        mlevel = maxd
    mlevel = max(mlevel, 1)
    nodes = [dict(ce=[cr[0], cr[1], cr[2], amax], chn=[0] * 8, volx=[0xFFFFFFFF, 0xFFFFFFFF], qval=[0.0, 0.0])]
    leafx = [0, 0]

    def overlap(ca, ea, cb, eb):
        for a in range(3):
            if _f(ca[a] - ea[a]) > _f(cb[a] + eb[a]) or _f(cb[a] - eb[a]) > _f(ca[a] + ea[a]):
                return False
        return True

    def build(x, c, e, c1, e1, level, bx):
        nodes[x]["ce"] = [c[0], c[1], c[2], e]
        nodes[x]["qval"] = [0.0, 0.0]
        if level >= mlevel - 1:
            if nodes[x]["volx"][bx] == 0xFFFFFFFF:
                nodes[x]["volx"][bx] = leafx[bx]
                leafx[bx] += 1
            return
        ee = _f(_f(0.5) * e)
        for i in range(8):
            sv = (_f(2.0 * (i % 2) - 1.0), _f(2.0 * ((i // 2) % 2) - 1.0), _f(2.0 * (i // 4) - 1.0))
            cc = [_f(c[a] + _f(sv[a] * ee)) for a in range(3)]
            ex = _f(_f(1.1) * ee)
            if not overlap(cc, [ex] * 3, c1, e1):
                continue
            if nodes[x]["chn"][i] == 0:
                nodes.append(dict(ce=None, chn=[0] * 8, volx=[0xFFFFFFFF, 0xFFFFFFFF], qval=[0.0, 0.0]))
                nodes[x]["chn"][i] = len(nodes) - 1
            build(nodes[x]["chn"][i], cc, ee, c1, e1, level + 1, bx)

    for bx in range(2):
        for c1, e1 in boxes_b[bx]:
            build(0, cr, amax, c1, e1, 0, bx)
    lvoxc = int(lvc)
    bricks = np.zeros((leafx[1], lvoxc ** 3), np.float32)
    sbricks = np.full((leafx[0], lvoxc ** 3), band, np.float32)   # the unresampled last node: outside
    inv_s = 1.0 / float(s)

    def box_sample(p, src=None, bnd=None):
        src = grid if src is None else src
        q = [float(p[a]) * inv_s for a in range(3)]
        i0 = [math.floor(v) for v in q]
        uvw = [q[a] - i0[a] for a in range(3)]

        def g(x, y, z):
            if x < 0 or y < 0 or z < 0 or x >= nx or y >= ny or z >= nz:
                return bnd if bnd is not None else _f(0.0)
            v = src[z, y, x]
            return max(min(v, bnd), _f(-bnd)) if bnd is not None else v

        def lp(a, b, w):
            return _f(a + _f(float(_f(b - a)) * w))
        d_ = [[[g(i0[0] + i, i0[1] + j, i0[2] + k) for k in range(2)] for j in range(2)] for i in range(2)]
        return lp(lp(lp(d_[0][0][0], d_[0][0][1], uvw[2]), lp(d_[0][1][0], d_[0][1][1], uvw[2]), uvw[1]),
                  lp(lp(d_[1][0][0], d_[1][0][1], uvw[2]), lp(d_[1][1][0], d_[1][1][1], uvw[2]), uvw[1]), uvw[0])

    nv = _f(lvoxc)
    NB = 0xFFFFFFFF
    for i in range(len(nodes) - 1):      # the reference resamples nodes [0, index) only
        n = nodes[i]
        if n["volx"][0] == NB:
            if n["volx"][1] == NB:
                continue
            if sdf is not None and box_sample(n["ce"][:3], sdf, band) < 0.0:
                n["volx"][1] = NB         # fog leaf inside the surface: dropped
                continue
        n["qval"] = [float(np.finfo(np.float32).max), 0.01]
        qf = _f(0.01)
        qs = _f(np.finfo(np.float32).max)
        for j in range(lvoxc ** 3):
            idx = (j % lvoxc, (j // lvoxc) % lvoxc, j // (lvoxc * lvoxc))
            p = []
            for a in range(3):
                nn = _f(_f(_f(nv - _f(_f(2.0) * _f(idx[a]))) - _f(1.0)) / _f(nv - _f(1.0)))
                p.append(_f(n["ce"][a] - _f(n["ce"][3] * nn)))
            if n["volx"][0] != NB:
                v = box_sample(p, sdf, band)
                sbricks[n["volx"][0], j] = v
                qs = min(qs, v)
            if n["volx"][1] != NB:
                v = max(min(box_sample(p), _f(1.0)), _f(0.0))
                bricks[n["volx"][1], j] = v
                qf = max(qf, v)
        n["qval"] = [float(qs), float(qf)]
    if sdf is not None:
        return nodes, bricks, lvoxc, mlevel, sbricks
    return nodes, bricks, lvoxc, mlevel


def map_env(texels, rd):
    """MapEnv::Evaluate then pow 2.2 (src/KernelSampler.cpp:384-421, src/kernel.cpp:362-364)
    for directions rd (n, 3); texels (h, w, 4)."""
    tex = f32(texels)
    h, w = tex.shape[:2]
    rd = f32(rd).reshape(-1, 3)
    out = np.zeros_like(rd)
    lib().or_map_env(fp(tex), C.c_uint32(w), C.c_uint32(h), fp(rd), C.c_uint32(len(rd)), fp(out))
    return out

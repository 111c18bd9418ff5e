"""Pins of the CPU oracle's point functions (no GPU needed).

Each test checks the oracle against something it did not compute itself:
published known answers, the reference's own sky-model source compiled from
where it lies (oracle/_ref), closed forms, brute force, or independently
measured properties of the reference's tables.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ------------------------------------------------------------------ RNG ----
def _kat():
    rows = []
    for line in open(os.path.join(ROOT, "tests", "golden", "philox_kat.txt")):
        a, b = line.split("->")
        v = [int(x, 16) for x in a.split()]
        rows.append((v[:4], v[4:], [int(x, 16) for x in b.split()]))
    return rows


def test_philox_known_answers(oracle):
    """Random123 / SC'11 known answers, regenerated with rocRAND's independent
    implementation (tests/golden/philox_kat_rocrand.cpp)."""
    rows = _kat()
    assert rows[0][2] == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    ctr = np.array([r[0] for r in rows], np.uint32)
    key = np.array([r[1] for r in rows], np.uint32)
    out = oracle.philox(ctr, key)
    assert np.array_equal(out, np.array([r[2] for r in rows], np.uint32))


def test_ray_stream_xoshiro_first_outputs(oracle):
    """xoshiro128** from a known state: hand-derived first outputs for the
    state (1,2,3,4) are 11520 and 0; the stream seeds its state from Philox."""
    rows = _kat()
    # seed state = philox((sample, ray, 0, 0), (key0, key1)); pick the KAT with
    # counter (1,2,3,4)? Not applicable; check through the uniforms instead:
    st = oracle.philox(np.array([[5, 9, 0, 0]], np.uint32), np.array([[17, 23]], np.uint32))[0]
    s = [int(x) for x in st]

    def nxt(s):
        r = (((s[1] * 5) & 0xFFFFFFFF) << 7 | ((s[1] * 5) & 0xFFFFFFFF) >> 25) & 0xFFFFFFFF
        r = (r * 9) & 0xFFFFFFFF
        t = (s[1] << 9) & 0xFFFFFFFF
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t
        s[3] = ((s[3] << 11) | (s[3] >> 21)) & 0xFFFFFFFF
        return r
    t = [1, 2, 3, 4]
    assert [nxt(t), nxt(t)] == [11520, 0]
    expect = [(((nxt(s) >> 8) | 1) * 2.0 ** -24) for _ in range(16)]
    got = oracle.ray_stream([17], [5], [9], 23, 16)[0]
    assert np.array_equal(got, np.float32(expect))
    assert rows  # KAT file present


def test_uniforms_open_interval(oracle):
    u = oracle.ray_stream(np.arange(64), np.zeros(64), np.arange(64), 99, 512).reshape(-1)
    assert u.min() > 0.0 and u.max() < 1.0
    k = u.astype(np.float64) * 2 ** 24
    assert np.all(k == np.round(k)) and np.all(np.round(k) % 2 == 1)
    assert abs(u.mean() - 0.5) < 0.01


# ----------------------------------------------------------------- Mie -----
def test_mie_table_properties():
    """Properties of the reference's table measured independently: the
    EvaluateRGB lerp integrates to ~1.043 over the sphere with mean cosine
    ~0.750 (bin 0 = backscatter, so a flipped index gives a negative mean)."""
    from oracle import fp64
    ct = np.linspace(-1, 1, 200001)
    p = fp64.mie_eval(ct)
    integ = np.trapezoid(p, ct, axis=0) * 2 * np.pi
    mc = np.trapezoid(p * ct[:, None], ct, axis=0) * 2 * np.pi / integ
    assert np.all(np.abs(integ - 1.043) < 0.003), integ
    assert np.all(np.abs(mc - 0.750) < 0.004), mc


def test_mie_eval_fp32_vs_fp64(oracle):
    from oracle import fp64
    rng = np.random.default_rng(3)
    ct = np.concatenate([rng.uniform(-1, 1, 4000), [-1.0, 1.0, 0.0, 1.5, -1.5]]).astype(np.float32)
    got = oracle.mie_eval(ct).astype(np.float64)
    ref = fp64.mie_eval(ct.astype(np.float64))
    rel = np.abs(got - ref) / np.abs(ref)
    # float32 acos and lerp: well inside 1e-5 except where the bin edge moves
    assert np.quantile(rel, 0.999) < 1e-5, np.quantile(rel, [0.5, 0.999])
    # table nodes: ct = -cos(pi*i/256) returns pdf[i] (bins 1..254)
    i = np.arange(1, 255)
    node = oracle.mie_eval((-np.cos(np.pi * i / 256.0)).astype(np.float32))
    assert np.allclose(node, oracle.MIE_PDF[i], rtol=2e-4)


def test_mie_sample_rule_and_monotone_keys(oracle):
    """MiePhase::Sample's quantised scan: the oracle's linear scan equals an
    independent vectorised evaluation over a dense grid of the generator's
    u values, the keys are non-decreasing (which makes the kernels' binary
    search equivalent), and the sampled mean cosine is ~0.737."""
    from oracle import fp64
    u = ((np.arange(0, 1 << 23, 97, dtype=np.int64) * 2 + 1) * 2.0 ** -24).astype(np.float32)
    th_scan = oracle.mie_theta(u[::41])
    got_vec, _, _ = fp64.mie_theta_fp32(u[::41])
    assert np.array_equal(th_scan, got_vec)
    for lo in range(0, len(u), 20000):
        th, keys, m = fp64.mie_theta_fp32(u[lo:lo + 20000])
        assert np.all(np.diff(keys, axis=1) >= 0)
        # the set of keys below u is a prefix
        assert np.all(np.diff(m.astype(np.int8), axis=1) <= 0)
    th, _, _ = fp64.mie_theta_fp32(u)
    ct = -np.cos(np.pi * th.astype(np.float64))
    assert abs(ct.mean() - 0.737) < 0.003, ct.mean()


def test_mie_sample_direction_frame(oracle):
    rng = np.random.default_rng(5)
    iv = rng.normal(size=(500, 3))
    iv /= np.linalg.norm(iv, axis=1, keepdims=True)
    u1, u2 = rng.random(500).astype(np.float32), rng.random(500).astype(np.float32)
    d = oracle.mie_sample(iv.astype(np.float32), u1, u2).astype(np.float64)
    th = oracle.mie_theta(u1).astype(np.float64) * np.pi
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-5)
    assert np.allclose((d * iv).sum(axis=1), -np.cos(th), atol=2e-5)


# ------------------------------------------------------------------ HG -----
@pytest.mark.parametrize("g", [0.75, -0.3, 0.2])
def test_hg_closed_forms(oracle, g):
    ct = np.linspace(-1, 1, 20001)
    p = oracle.hg_eval(g, ct).astype(np.float64)
    assert abs(np.trapezoid(p, ct) * 2 * np.pi - 1.0) < 1e-4
    assert abs(np.trapezoid(p * ct, ct) * 2 * np.pi - g) < 1e-4
    rng = np.random.default_rng(1)
    iv = np.tile(np.array([[0.6, 0.0, 0.8]], np.float32), (20000, 1))
    d = oracle.hg_sample(g, iv, rng.random(20000).astype(np.float32), rng.random(20000).astype(np.float32))
    cts = np.sort((d.astype(np.float64) @ iv[0].astype(np.float64)))
    # CDF of cos theta under HG
    cdf = (1 - g * g) / (2 * g) * (1 / np.sqrt(1 + g * g - 2 * g * cts) - 1 / (1 + g))
    ks = np.max(np.abs(cdf - np.arange(1, len(cts) + 1) / len(cts)))
    assert ks < 0.015, ks


# ----------------------------------------------------------------- sun -----
def test_sun_sampling_and_pdf(oracle):
    d = np.array([math.cos(math.radians(30)), 0.0, math.sin(math.radians(30))], np.float32)
    ang = 0.01
    rng = np.random.default_rng(2)
    n = 20000
    u1, u2 = rng.random(n).astype(np.float32), rng.random(n).astype(np.float32)
    probe = np.concatenate([np.tile(d, (n // 2, 1)), np.tile([0, 0, 1], (n - n // 2, 1))]).astype(np.float32)
    sd, pdf, ev = oracle.sun(d, [1, 1, 1], ang, u1, u2, probe)
    ct = sd.astype(np.float64) @ d.astype(np.float64)
    cosa = math.cos(ang)
    assert ct.min() > cosa - 1e-6
    v = np.sort((1 - ct) / (1 - cosa))     # uniform in [0,1] for a cone sampled uniform in cos
    assert np.max(np.abs(v - np.arange(1, n + 1) / n)) < 0.03
    # the reference's Pdf: 1/(2 pi (1 - cos^2)), not the cone's 1/(2 pi (1 - cos))
    assert np.allclose(pdf[: n // 2], 1.0 / (2 * np.pi * (1 - np.float32(cosa) ** 2)), rtol=1e-3)
    assert np.all(pdf[n // 2:] == 0) and np.all(ev[n // 2:] == 0) and np.all(ev[: n // 2] == 1)


# ----------------------------------------------------------------- sky -----
def _ref_hosek():
    path = os.path.join(ROOT, "oracle", "_ref", "libarhosek_ref.so")
    if not os.path.exists(path):
        pytest.skip("oracle/_ref not built (needs the reference tree; `make -C oracle ref`)")
    L = C.CDLL(path)
    L.arhosek_rgb_skymodelstate_alloc_init.restype = C.c_void_p
    L.arhosek_rgb_skymodelstate_alloc_init.argtypes = [C.c_double, C.c_double, C.c_double]
    L.arhosek_tristim_skymodel_radiance.restype = C.c_double
    L.arhosek_tristim_skymodel_radiance.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int]
    L.arhosekskymodelstate_free.argtypes = [C.c_void_p]
    return L


@pytest.mark.parametrize("T,a,el", [(3.0, 0.1, 0.5235988), (7.0, 0.15, 0.3), (1.0, 0.0, 1.2), (9.5, 1.0, 0.05),
                                    (10.0, 0.5, 0.7)])
def test_hosek_state_matches_reference_source(pkg, oracle, T, a, el):
    """The oracle's and the library's Hosek state against the reference's own
    ArHosekSkyModel.cpp compiled from the reference tree (oracle/_ref)."""
    L = _ref_hosek()
    st = L.arhosek_rgb_skymodelstate_alloc_init(T, a, el)
    raw = np.ctypeslib.as_array(C.cast(st, C.POINTER(C.c_double)), shape=(11 * 9 + 11,))
    ref_cfg = raw[:27].reshape(3, 9).copy()
    ref_rad = raw[99:102].copy()
    L.arhosekskymodelstate_free(st)
    cfg, rad = oracle.hosek_state(T, a, el)
    assert np.array_equal(cfg, ref_cfg) and np.array_equal(rad, ref_rad)
    c27, r3 = np.zeros(27), np.zeros(3)
    pkg._lib.check(pkg._lib.lib().drk_probe_hosek_state(T, a, el, pkg._lib.dptr(c27), pkg._lib.dptr(r3)))
    assert np.array_equal(c27.reshape(3, 9), ref_cfg) and np.array_equal(r3, ref_rad)


def test_sky_formula_matches_reference_radiance(oracle):
    """Where the kernel's clamps are inactive (rd.z >= 0, 0 <= cos gamma <= 1)
    its sky equals 1e-3 * (ArHosekSkyModel radiance)^2.2 with the reference's
    gamma construction; the FP64 restatement is checked against the reference
    source, and the FP32 oracle against the FP64 restatement."""
    from oracle import fp64
    L = _ref_hosek()
    d = np.array([math.cos(math.radians(30)), 0.0, math.sin(math.radians(30))], np.float32)
    el = oracle.sun_elevation(d)
    st = L.arhosek_rgb_skymodelstate_alloc_init(3.0, 0.1, el)
    rng = np.random.default_rng(4)
    rd = rng.normal(size=(3000, 3))
    rd[:, 2] = np.abs(rd[:, 2])
    rd /= np.linalg.norm(rd, axis=1, keepdims=True)
    cph = fp64.sky_cph(rd, d)
    keep = (cph >= 0) & (cph <= 1) & (np.abs(rd[:, 1]) > 1e-3)
    rd = rd[keep]
    cph = cph[keep]
    ref = np.array([[1e-3 * L.arhosek_tristim_skymodel_radiance(st, math.acos(z), math.acos(c), ch) ** 2.2
                     for ch in range(3)] for z, c in zip(rd[:, 2], cph)])
    L.arhosekskymodelstate_free(st)
    cfg, rad = oracle.hosek_state(3.0, 0.1, el)
    f64 = fp64.sky(d, 3.0, 0.1, rd, cfg=cfg, rad=rad)
    assert np.allclose(f64, ref, rtol=1e-12)
    f32 = oracle.sky(d, 3.0, 0.1, rd.astype(np.float32)).astype(np.float64)
    full = fp64.sky(d, 3.0, 0.1, rd.astype(np.float32).astype(np.float64))
    rel = np.abs(f32 - full) / np.abs(full)
    assert np.quantile(rel, 0.99) < 1e-5 and rel.max() < 1e-4, np.quantile(rel, [0.5, 0.99, 1.0])


# --------------------------------------------------------- trilinear -------
def test_trilinear_vertices_and_linearity(oracle):
    from oracle import fp64
    rng = np.random.default_rng(8)
    brick = rng.random(512).astype(np.float32)
    idx = np.stack(np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij"), -1).reshape(-1, 3)[:, ::-1]
    # a node whose vertex coordinates and index arithmetic are exact in float32
    ce_exact = np.array([1.0, 1.0, 1.0, 0.875], np.float32)
    p = (0.125 + 0.25 * idx).astype(np.float32)
    got = oracle.voxel(brick, p, ce_exact, 8)
    assert np.array_equal(got, brick[idx[:, 2] * 64 + idx[:, 1] * 8 + idx[:, 0]])
    ce = np.array([0.3, -0.2, 0.55, 0.125], np.float32)
    q = ce[:3] - ce[3] + rng.random((3000, 3)) * 2 * ce[3]
    got = oracle.voxel(brick, q, ce, 8).astype(np.float64)
    assert np.allclose(got, fp64.trilinear(brick, q, ce, 8), atol=2e-6)
    lin = (0.25 + 0.5 * (idx[:, 0] / 7.0)).astype(np.float32)
    b2 = np.zeros(512, np.float32)
    b2[idx[:, 2] * 64 + idx[:, 1] * 8 + idx[:, 0]] = lin
    got = oracle.voxel(b2, q, ce, 8).astype(np.float64)
    assert np.allclose(got, 0.25 + 0.5 * (q[:, 0] - (ce[0] - ce[3])) / (2 * ce[3]), atol=2e-6)


# ----------------------------------------------------------- traversal -----
def brute_leaves(ob, ro, rd):
    """Slab test of every leaf in float64, ordered by entry distance."""
    leaves = np.nonzero((ob["volx"][:, 0] != 0xFFFFFFFF) | (ob["volx"][:, 1] != 0xFFFFFFFF))[0]
    c = ob["ce"][leaves, :3].astype(np.float64)
    e = ob["ce"][leaves, 3:4].astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        ta = (c - e - ro) / rd
        tb = (c + e - ro) / rd
    t0 = np.minimum(ta, tb).max(axis=1)
    t1 = np.maximum(ta, tb).min(axis=1)
    hit = (t0 < t1) & (t1 >= 0)
    o = np.argsort(t0[hit])
    return leaves[hit][o], t0[hit][o], t1[hit][o]


def test_traversal_matches_brute_force(pkg, oracle):
    g = pkg.configs.synth_spheres(64, 9, spheres=10)
    sc = pkg.Scene.from_grid(g)
    rng = np.random.default_rng(6)
    n = 600
    ro = np.concatenate([rng.uniform(-1.5, 2.5, (n // 2, 3)), rng.uniform(0.1, 0.9, (n - n // 2, 3))])
    rd = rng.normal(size=(n, 3))
    rd /= np.linalg.norm(rd, axis=1, keepdims=True)
    nodes, t0, t1, cnt = oracle.traverse(sc.ob, ro.astype(np.float32), rd.astype(np.float32))
    checked = 0
    for i in range(n):
        bn, b0, b1 = brute_leaves(sc.ob, ro[i].astype(np.float32).astype(np.float64),
                                  rd[i].astype(np.float32).astype(np.float64))
        if len(bn) and np.min(np.diff(b0)) < 1e-6 if len(bn) > 1 else False:
            continue  # ties in entry order: ambiguous for the brute force
        assert cnt[i] == len(bn), (i, cnt[i], len(bn))
        assert np.array_equal(nodes[i, :cnt[i]], bn), i
        assert np.allclose(t0[i, :cnt[i]], b0, atol=1e-5) and np.allclose(t1[i, :cnt[i]], b1, atol=1e-5)
        checked += 1
    assert checked > n * 0.9


# --------------------------------------------------------- transmittance ---
def optical_depth(sc, ro, rd, nodes, t0, t1, k, steps=256):
    """FP64 midpoint integral of the trilinear brick density along the ray's
    leaf segments (the last node is transparent: never resampled)."""
    from oracle import fp64
    tau = 0.0
    last = sc.node_count - 1
    bricks = sc.pvol[1].reshape(-1, sc.lvoxc3)
    for j in range(k):
        n = int(nodes[j])
        if n == last or t1[j] <= t0[j]:
            continue
        a, b = max(float(t0[j]), 0.0), float(t1[j])
        if b <= a:
            continue
        ts = a + (np.arange(steps) + 0.5) * (b - a) / steps
        pts = ro[None, :].astype(np.float64) + ts[:, None] * rd[None, :].astype(np.float64)
        rho = fp64.trilinear(bricks[sc.ob["volx"][n, 1]], pts, sc.ob["ce"][n], sc.lvoxc)
        tau += rho.sum() * (b - a) / steps
    return tau


def test_homogeneous_cube_alpha_is_beer_lambert(pkg, oracle):
    """Camera paths scatter with probability 1 - exp(-sigma_t * tau), tau the
    optical depth of the trilinear brick density along the ray (the bricks
    of a constant cube fall off at its faces, where the resampling meets the
    zero background). The last leaf never gets resampled by the reference
    (majorant 0), so it is transparent."""
    n = 32
    g = np.ones((n, n, n), np.float32)
    wl = pkg.configs.Workload("cube", g, pkg.Scene.from_grid(g), 16, 16, 512, 0, 3.0)
    r = oracle.render(wl, seed=5)
    alpha = r[0][..., 3].astype(np.float64) / wl.spp
    ys, xs = np.mgrid[0:16, 0:16]
    ro, rd, _ = oracle.camera_rays(wl.viewi, wl.proji, 16, 16, xs.ravel(), ys.ravel())
    nodes, t0, t1, cnt = oracle.traverse(wl.scene.ob, ro, rd)
    p = np.zeros(len(ro))
    for i in range(len(ro)):
        p[i] = 1 - math.exp(-3.0 * optical_depth(wl.scene, ro[i], rd[i], nodes[i], t0[i], t1[i], cnt[i]))
    sig = np.sqrt(np.maximum(p * (1 - p), 1e-12) / wl.spp)
    z = (alpha.ravel() - p) / np.maximum(sig, 1e-9)
    assert np.abs(z).max() < 5.0
    assert abs(z.mean()) < 4.0 / math.sqrt(len(z))
    assert (cnt > 0).mean() > 0.3


def test_fast_oracle_build_computes_the_same_numbers(pkg, oracle):
    """The -O3 x86-64-v3 build of oracle.c (bench.py's second CPU-baseline
    build) renders exactly what the -O2 build renders (no contraction)."""
    wl = pkg.configs.tiny(w=24, h=16, spp=8, scattevs=4)
    a = oracle.render(wl, seed=3)
    b = oracle.render(wl, seed=3, fast=True)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)

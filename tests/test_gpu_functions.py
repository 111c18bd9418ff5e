"""Function-level parity of the gfx950 device code against the CPU oracle.

Each test runs one device function through the C-ABI probe hooks. Integer and
index work (RNG, traversal, trilinear lookup, camera rays) must be bit-exact;
floating-point point functions (phase, sky) are held to 1e-5 relative against
the FP64 restatement, as the accuracy bar for FP32 requires.
"""
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SUN = np.array([math.cos(math.radians(30)), 0.0, math.sin(math.radians(30))], np.float32)


def test_philox_bit_exact(pkg, oracle, gpu):
    rows = []
    for line in open(os.path.join(ROOT, "tests", "golden", "philox_kat.txt")):
        a, b = line.split("->")
        v = [int(x, 16) for x in a.split()]
        rows.append((v[:4], v[4:], [int(x, 16) for x in b.split()]))
    ctr = np.array([r[0] for r in rows], np.uint32)
    key = np.array([r[1] for r in rows], np.uint32)
    assert np.array_equal(pkg.probe.philox(ctr, key), np.array([r[2] for r in rows], np.uint32))
    rng = np.random.default_rng(0)
    ctr = rng.integers(0, 2 ** 32, (4096, 4), dtype=np.uint64).astype(np.uint32)
    key = rng.integers(0, 2 ** 32, (4096, 2), dtype=np.uint64).astype(np.uint32)
    assert np.array_equal(pkg.probe.philox(ctr, key), oracle.philox(ctr, key))


def test_ray_streams_bit_exact(pkg, oracle, gpu):
    rng = np.random.default_rng(1)
    n = 2048
    k0 = rng.integers(0, 2 ** 20, n).astype(np.uint32)
    s = rng.integers(0, 4096, n).astype(np.uint32)
    r = rng.integers(0, 129, n).astype(np.uint32)
    assert np.array_equal(pkg.probe.ray_stream(k0, s, r, 12345, 64), oracle.ray_stream(k0, s, r, 12345, 64))


def test_free_flight_log_equals_logf_for_every_uniform(pkg, gpu):
    """The tracker's log (hardware log2 scaled by a two-part ln 2, without
    the library's denormal and non-finite handling) equals the device
    library's logf bit for bit on all 2^23 uniforms the RNG can return, and
    both are within two ulps of the float64 logarithm (measured: 1.88, next
    to u = 1 where the result is tiny)."""
    a, b = pkg.probe.log_unit()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    u = ((np.arange(1 << 23, dtype=np.float64) * 2 + 1) * 2.0 ** -24)
    exact = np.log(u)
    ulp = np.spacing(np.abs(exact).astype(np.float32)).astype(np.float64)
    assert np.max(np.abs(a.astype(np.float64) - exact) / ulp) <= 2.0


def test_camera_rays_bit_exact(pkg, oracle, gpu):
    wl = pkg.configs.tiny(w=96, h=64)
    ys, xs = np.mgrid[0:64, 0:96]
    g = pkg.probe.camera(wl.viewi, wl.proji, 96, 64, xs.ravel(), ys.ravel())
    c = oracle.camera_rays(wl.viewi, wl.proji, 96, 64, xs.ravel(), ys.ravel())
    for a, b in zip(g, c):
        assert np.array_equal(a, b)


def test_mie_eval_within_1e5(pkg, oracle, gpu):
    from oracle import fp64
    rng = np.random.default_rng(2)
    ct = np.concatenate([rng.uniform(-1, 1, 20000), [-1, 1, 0, 1.2, -1.2]]).astype(np.float32)
    g = pkg.probe.mie_eval(ct).astype(np.float64)
    ref = fp64.mie_eval(ct.astype(np.float64))
    rel = np.abs(g - ref) / np.abs(ref)
    assert np.quantile(rel, 0.999) < 1e-5 and rel.max() < 1e-3, np.quantile(rel, [0.5, 0.999, 1])
    c = oracle.mie_eval(ct).astype(np.float64)
    assert np.quantile(np.abs(g - c) / np.abs(c), 0.999) < 1e-5


def test_mie_sample_same_bin_as_reference_scan(pkg, oracle, gpu):
    """The kernels' binary search picks the same quantised angle as the
    reference's linear scan for every u."""
    rng = np.random.default_rng(3)
    m = rng.integers(0, 1 << 23, 50000) * 2 + 1           # odd multiples of 2^-24
    # and, for the per-bucket search bounds, every bucket's first and last
    # uniform and the neighbourhood of the two places where the answer steps
    # back as u1 grows (u1 ~ 0.98223 and 0.98948)
    per = (1 << 24) // 1024
    edges = np.concatenate([np.arange(1024) * per + 1, (np.arange(1024) + 1) * per - 1])
    dips = np.concatenate([np.arange(-64, 66, 2) + int(round(x * 2 ** 24)) // 2 * 2 + 1 for x in (0.9822312, 0.9894795)])
    m = np.concatenate([m, edges, dips])
    n = len(m)
    u1 = (m * 2.0 ** -24).astype(np.float32)
    u2 = rng.random(n).astype(np.float32)
    iv = np.tile(np.array([[0.48, 0.6, 0.64]], np.float32), (n, 1))
    iv /= np.linalg.norm(iv, axis=1, keepdims=True)
    d = pkg.probe.mie_sample(iv, u1, u2).astype(np.float64)
    th = oracle.mie_theta(u1).astype(np.float64) * np.pi
    assert np.allclose((d * iv).sum(axis=1), -np.cos(th), atol=3e-6)
    c = oracle.mie_sample(iv[:4000], u1[:4000], u2[:4000]).astype(np.float64)
    assert np.allclose(d[:4000], c, atol=3e-6)


@pytest.mark.parametrize("g", [0.75, -0.4])
def test_hg_within_1e5(pkg, oracle, gpu, g):
    from oracle import fp64
    ct = np.linspace(-1, 1, 4001).astype(np.float32)
    e = pkg.probe.hg_eval(g, ct).astype(np.float64)
    assert np.max(np.abs(e - fp64.hg_eval(g, ct)) / fp64.hg_eval(g, ct)) < 1e-5
    rng = np.random.default_rng(4)
    iv = rng.normal(size=(2000, 3)).astype(np.float32)
    iv /= np.linalg.norm(iv, axis=1, keepdims=True)
    u1, u2 = rng.random(2000).astype(np.float32), rng.random(2000).astype(np.float32)
    assert np.allclose(pkg.probe.hg_sample(g, iv, u1, u2), oracle.hg_sample(g, iv, u1, u2), atol=3e-6)


def test_sun_matches_oracle(pkg, oracle, gpu):
    rng = np.random.default_rng(5)
    n = 4000
    u1, u2 = rng.random(n).astype(np.float32), rng.random(n).astype(np.float32)
    dirs = np.concatenate([np.tile(SUN, (n // 2, 1)), rng.normal(size=(n - n // 2, 3))]).astype(np.float32)
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    g = pkg.probe.sun(SUN, (1, 0.9, 0.8), 0.01, u1, u2, dirs)
    c = oracle.sun(SUN, (1, 0.9, 0.8), 0.01, u1, u2, dirs)
    assert np.allclose(g[0], c[0], atol=3e-6)
    assert np.array_equal(g[1], c[1]) and np.array_equal(g[2], c[2])


def test_sky_within_1e5_of_fp64(pkg, gpu):
    from oracle import fp64
    rng = np.random.default_rng(6)
    rd = rng.normal(size=(20000, 3))
    rd /= np.linalg.norm(rd, axis=1, keepdims=True)
    rd = rd[np.abs(rd[:, 1]) > 1e-3].astype(np.float32)
    g = pkg.probe.sky(SUN, 3.0, 0.1, rd).astype(np.float64)
    ref = fp64.sky(SUN, 3.0, 0.1, rd.astype(np.float64))
    ok = np.isfinite(ref).all(axis=1)
    rel = np.abs(g[ok] - ref[ok]) / np.abs(ref[ok])
    assert np.quantile(rel, 0.99) < 2e-6 and rel.max() < 1e-5, np.quantile(rel, [0.5, 0.99, 1])
    assert np.array_equal(np.isfinite(g), np.isfinite(ref))


@pytest.mark.parametrize("leaf_rcp", [False, True])
def test_trilinear_bit_exact(pkg, oracle, gpu, leaf_rcp):
    """The lookup with the division by the leaf extent, and with the
    reciprocal and Markstein correction the kernels use when all leaves share
    one extent: bit for bit the oracle's, also for points on and next to the
    cell planes (numerators 0 and near it, where the division is kept) and
    extents that are not powers of two."""
    rng = np.random.default_rng(7)
    for lvoxc in (8, 16):
        brick = rng.random(lvoxc ** 3).astype(np.float32)
        for ce in (np.array([0.41, 0.23, -0.6, 0.0625], np.float32),
                   np.array([0.0190, 0.0190, 0.5, 0.0190], np.float32),
                   np.array([1.3, -2.7, 0.77, 0.3109712], np.float32)):
            p = (ce[:3] - ce[3] * 1.01 + rng.random((5000, 3)) * 2.02 * ce[3]).astype(np.float32)
            # points on the lower faces (zero numerators), one ulp either side,
            # and on interior cell planes
            lo = (ce[:3] - ce[3]).astype(np.float32)
            edge = p[:300].copy()
            edge[:100, 0] = lo[0]
            edge[100:200, 1] = np.nextafter(lo[1], np.float32(np.inf))
            edge[200:300, 2] = np.nextafter(lo[2], np.float32(-np.inf))
            planes = (lo[None, :] + (2 * ce[3] / (lvoxc - 1)) * rng.integers(0, lvoxc, (300, 3))).astype(np.float32)
            p = np.concatenate([p, edge, planes]).astype(np.float32)
            got = pkg.probe.trilinear(brick, lvoxc, ce, p, leaf_rcp=leaf_rcp)
            assert np.array_equal(got.view(np.uint32), oracle.voxel(brick, p, ce, lvoxc).view(np.uint32))


@pytest.mark.parametrize("walk", ["octree", "grid"])
@pytest.mark.parametrize("cfg", ["c0", "spheres64", "c1"])
def test_traversal_bit_exact(pkg, oracle, gpu, cfg, walk):
    """Leaf sequences (node indices and [tr0, tr1]) of camera rays and of
    secondary rays starting inside the volume, against the oracle's recursive
    full traverser, for both walks of the kernels (the Revelles stack walk
    and the leaf-grid walk with its hand-overs)."""
    if cfg == "c0":
        wl = pkg.configs.c0(w=64, h=64)
        sc = wl.scene
        ys, xs = np.mgrid[0:64, 0:64]
        ro, rd, _ = oracle.camera_rays(wl.viewi, wl.proji, 64, 64, xs.ravel(), ys.ravel())
    else:
        sc = (pkg.Scene.from_grid(pkg.configs.synth_spheres(64, 3, spheres=12)) if cfg == "spheres64"
              else pkg.configs.c1(w=8, h=8).scene)
        ro = np.zeros((0, 3), np.float32)
        rd = np.zeros((0, 3), np.float32)
    sc.set_walk(pkg._lib.DRK_WALK_GRID if walk == "grid" else pkg._lib.DRK_WALK_OCTREE)
    assert sc.walk_info()[2] == (walk == "grid")
    rng = np.random.default_rng(8)
    m = 20000
    ro2 = rng.uniform(0.0, 1.0, (m, 3)).astype(np.float32)
    rd2 = rng.normal(size=(m, 3))
    rd2 /= np.linalg.norm(rd2, axis=1, keepdims=True)
    # axis-aligned and diagonal directions exercise the tie rules; origins on
    # the cell boundaries exercise the hand-over to the Revelles walk
    rd2[:200] = np.array([[1, 0, 0], [0, -1, 0], [0, 0, 1], [-1, 0, 0]] * 50)
    rd2[200:300] = np.array([1, 1, 1]) / math.sqrt(3)
    rd2[300:600] = rng.choice([-1.0, 1.0], (300, 3)) / math.sqrt(3)
    c = sc.ob["ce"][0]
    ro2[300:900] = (c[:3] + rng.integers(-16, 17, (600, 3)) / 16.0 * c[3]).astype(np.float32)
    ro = np.concatenate([ro, ro2]).astype(np.float32)
    rd = np.concatenate([rd, rd2]).astype(np.float32)
    g = pkg.probe.traverse(sc, ro, rd)
    c = oracle.traverse(sc.ob, ro, rd)
    assert np.array_equal(g[3], c[3])
    for a, b in zip(g[:3], c[:3]):
        assert np.array_equal(a, b)
    assert g[3].max() > 3

"""The oracle's packet path (no GPU needed).

or_render_packets traces the estimator in the reference's CPU structure --
2x2 pixel packets in four lanes that run the leaf loop, the Woodcock loop and
both recursive calls in lockstep under masks (src/kernel.cpp:128-487,
src/smmath.inl:4-6). Each lane draws from its own stream only while it is
active and rounds as the one-lane restatement does, so its sums (both
buffers and both sums of squares) must equal or_render's bit for bit on
every branch of the estimator: Mie/HG, Hosek/MapEnv/NullEnv, SDF leaves,
depth compositing, odd and overhanging tiles, and a tile of the 256^3 frame.
It is the CPU baseline's packet build.
"""
import numpy as np
import pytest

import radiance_cases as RC


def _same_bits(a, b):
    for x, y in zip(a, b):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


def _both(oracle, wl, fast=False, **kw):
    a = oracle.render(wl, fast=fast, **kw)
    b = oracle.render(wl, fast=fast, packets=True, **kw)
    return a, b


@pytest.mark.parametrize("fast", [False, True])
def test_packets_equal_one_lane_on_c0(pkg, oracle, fast):
    wl = pkg.configs.c0(w=48, h=40, spp=4)
    a, b = _both(oracle, wl, fast=fast, seed=3)
    _same_bits(a, b)
    assert a[0][..., 3].sum() > 100 and a[1][..., :3].sum() > 0


def test_packets_equal_one_lane_with_hg_mapenv_nullenv(pkg, oracle):
    wl = pkg.configs.tiny(w=24, h=24, spp=6, scattevs=5)
    wl.phase = pkg.HGPhase(0.6)
    th, ph = np.mgrid[0:16, 0:32]
    tex = np.stack([0.3 + 0.2 * np.cos(np.pi * th / 15), 0.4 + 0.1 * np.sin(2 * np.pi * ph / 31),
                    0.6 + 0.0 * th, np.ones_like(th)], -1)
    wl.env = pkg.MapEnv(tex.astype(np.float32))
    a, b = _both(oracle, wl, seed=8)
    _same_bits(a, b)
    assert a[1][..., :3].sum() > 0
    wl.env = pkg.NullEnv()
    a, b = _both(oracle, wl, seed=9)
    _same_bits(a, b)


def test_packets_equal_one_lane_with_sdf_and_fog(pkg, oracle):
    from test_octree_build import sphere_sdf
    n = 40
    sdf = sphere_sdf(n, r=0.25)
    fog = np.zeros((n, n, n), np.float32)
    fog[8:32, 8:32, 8:32] = 0.6
    sc = pkg.Scene.from_grid(fog, sdf=sdf)
    wl = pkg.configs.Workload("sdf_fog", sdf, sc, 20, 20, 8, 3, 6.0)
    a, b = _both(oracle, wl, seed=12)
    _same_bits(a, b)
    assert a[0][..., 3].sum() > 0


def test_packets_equal_one_lane_with_depth_and_odd_tiles(pkg, oracle):
    wl = pkg.configs.tiny(w=30, h=26, spp=5)
    depth = np.full((wl.h, wl.w), 1e30, np.float32)
    depth[:, : wl.w // 2] = 1.6
    a, b = _both(oracle, wl, seed=4, depth=depth)
    _same_bits(a, b)
    for x0, y0, tw, th in [(3, 5, 9, 7), (25, 21, 8, 8), (0, 0, 31, 27)]:
        a, b = _both(oracle, wl, seed=6, x0=x0, y0=y0, tilex=tw, tiley=th)
        _same_bits(a, b)


def test_packets_equal_one_lane_in_the_furnace(pkg, oracle):
    wl = RC.furnace_workload(pkg)
    a, b = _both(oracle, wl, seed=31, x0=8, y0=8, tilex=8, tiley=8, samples=64)
    _same_bits(a, b)


def test_packets_equal_one_lane_on_a_c1_tile(pkg, oracle):
    wl = pkg.configs.c1(spp=2)
    a, b = _both(oracle, wl, fast=True, seed=1, x0=480, y0=256, tilex=24, tiley=24)
    _same_bits(a, b)
    assert a[0][..., 3].sum() > 0

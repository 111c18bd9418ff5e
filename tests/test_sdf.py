"""SDF leaves (src/kernel.cpp:245-315, src/scene.cpp:874-904): a leaf with an
SDF brick has majorant 1, is entered at its entry point only when that point
lies ahead and outside the surface, and a tentative collision inside the
surface (distance <= 0) is a real one; outside, the leaf's fog (if any) decides.

CPU: for a solid sphere (no fog) those rules make the camera ray's collision a
Poisson event of rate sigma_t over the part of its path that the tracker
counts -- inside the surface within SDF leaves, plus the stretch before an SDF
leaf whose entry point is inside the surface -- so the oracle's alpha must be
1 - exp(-sigma_t * that length), computed here independently (FP64 distance
lookups along each ray). GPU: same-key agreement with the oracle on a sphere
with fog inside and around it, and the adjoint refusing SDF scenes."""
import math

import numpy as np
import pytest

from test_octree_build import sphere_sdf

NO = 0xFFFFFFFF


def _sphere_workload(pkg, n=48, r=0.3, w=24, h=24, spp=512, scattevs=0, sigma_t=8.0, fog=None):
    sdf = sphere_sdf(n, r=r)
    sc = pkg.Scene.from_grid(fog, sdf=sdf)
    return pkg.configs.Workload("sdf_sphere%d" % n, sdf, sc, w, h, spp, scattevs, sigma_t)


def _counted_length(oracle, sc, ro, rd, nodes, t0, t1, cnt, steps=400):
    from oracle import fp64
    sb = sc.pvol[0].reshape(sc.leafx[0], -1)
    td, meas = 0.0, 0.0
    for j in range(int(cnt)):
        n = int(nodes[j])
        vs = int(sc.ob["volx"][n, 0])
        if vs == NO:
            continue                      # no fog in this scene: a fog-only leaf counts nothing
        tra, trb = float(t0[j]), float(t1[j])
        ce = sc.ob["ce"][n]
        skip = False
        if tra > td:
            p0 = (ro + rd * np.float32(tra)).astype(np.float32)
            skip = oracle.voxel(sb[vs], p0, ce, sc.lvoxc) > 0.0
        if skip:
            a = tra
        else:
            meas += max(tra - td, 0.0)    # tentative points before the entry count as collisions
            a = max(td, tra)
        b = trb
        if b > a:
            t = a + (np.arange(steps) + 0.5) * (b - a) / steps
            pts = ro[None, :].astype(np.float64) + t[:, None] * rd[None, :].astype(np.float64)
            d = fp64.trilinear(sb[vs], pts, ce, sc.lvoxc)
            meas += float(np.mean(d <= 0.0)) * (b - a)
        td = trb
    return meas


def test_solid_sphere_alpha_is_poisson_over_the_counted_length(pkg, oracle):
    wl = _sphere_workload(pkg)
    r = oracle.render(wl, seed=17)
    alpha = r[0][..., 3].astype(np.float64) / wl.spp
    ys, xs = np.mgrid[0:wl.h, 0:wl.w]
    ro, rd, gm = oracle.camera_rays(wl.viewi, wl.proji, wl.w, wl.h, xs.ravel(), ys.ravel())
    nodes, t0, t1, cnt = oracle.traverse(wl.scene.ob, ro, rd)
    p = np.array([1.0 - math.exp(-wl.sigma_t * _counted_length(oracle, wl.scene, ro[i], rd[i], nodes[i], t0[i],
                                                                    t1[i], cnt[i])) for i in range(len(ro))])
    sig = np.sqrt(np.maximum(p * (1 - p), 1e-12) / wl.spp)
    z = (alpha.ravel() - p) / np.maximum(sig, 1e-9)
    assert ((p > 0.05) & (p < 0.95)).sum() >= 10          # the test sees partial probabilities
    assert (p > 0.99).sum() >= 10                          # and rays through the solid interior
    assert np.abs(z).max() < 5.0, np.abs(z).max()
    assert abs(z.mean()) < 4.0 / math.sqrt(len(z))


def test_no_fog_leaves_inside_and_band_background(pkg):
    wl = _sphere_workload(pkg, fog=np.ones((48, 48, 48), np.float32))
    sc = wl.scene
    fog_only = (sc.ob["volx"][:, 0] == NO) & (sc.ob["volx"][:, 1] != NO)
    c = sc.ob["ce"][fog_only][:, :3]
    d = np.linalg.norm(c - 0.5, axis=1) - 0.3
    assert (d > -0.05).all()     # no fog-only leaf is left deep inside the surface


@pytest.mark.gpu
def test_gpu_matches_oracle_with_sdf_and_fog(pkg, oracle, gpu):
    from test_gpu_render import render_gpu, same_key_agreement
    n = 48
    fog = pkg.configs.synth_fbm(n, 9, octaves=3, freq=3.0, threshold=0.15)
    wl = _sphere_workload(pkg, n=n, r=0.25, w=40, h=32, spp=16, scattevs=6, sigma_t=10.0, fog=fog)
    assert wl.scene.leafx[0] > 0 and wl.scene.leafx[1] > 0
    g = render_gpu(wl, 5)
    c = oracle.render(wl, seed=5)
    same_key_agreement(g, c)
    assert g[0][..., 3].sum() > 0


@pytest.mark.gpu
def test_adjoint_refuses_sdf_scenes(pkg, gpu):
    from droplet_render_amd import _lib as L
    wl = _sphere_workload(pkg, n=32, w=16, h=16, spp=2, fog=np.ones((32, 32, 32), np.float32) * 0.2)
    k = wl.kernel()
    z = np.zeros((16, 16, 4), np.float32)
    with pytest.raises(L.DrkError):
        k.adjoint(0, 0, 16, 16, 2, 1, 0, z, z)
    k.Destroy()

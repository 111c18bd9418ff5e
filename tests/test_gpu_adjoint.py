"""The gfx950 adjoint kernel: parity with the CPU oracle's adjoint (same RNG
keys), consistency with central finite differences of the GPU render (common
random numbers, majorants fixed), and the torch.autograd wrapper."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def raised_scene(pkg, wl, factor=1.25):
    sc = pkg.Scene(wl.scene.ob.copy(), wl.scene.pvol[1].copy(), wl.scene.lvoxc, wl.scene.mlevel)
    sc.raise_majorants(factor)
    return sc


def test_adjoint_matches_oracle(pkg, oracle, gpu):
    wl = pkg.configs.tiny(n=32, w=32, h=32, spp=32, scattevs=4)
    rng = np.random.default_rng(2)
    gl = rng.normal(size=(32, 32, 4)).astype(np.float32)
    gl[..., :3] *= 1000.0
    ge = rng.normal(size=(32, 32, 4)).astype(np.float32)
    k = wl.kernel()
    g = k.adjoint(0, 0, 32, 32, wl.spp, 5, 0, gl, ge).astype(np.float64)
    k.Destroy()
    c = oracle.adjoint(wl, 5, gl, ge).astype(np.float64)
    assert np.isfinite(g).all()
    cos = g @ c / (np.linalg.norm(g) * np.linalg.norm(c))
    rel = np.linalg.norm(g - c) / np.linalg.norm(c)
    assert cos > 0.999 and rel < 0.03, (cos, rel)
    assert np.array_equal(g == 0, c == 0) or np.mean((g == 0) != (c == 0)) < 0.01


def test_adjoint_device_accumulates(pkg, gpu):
    import torch
    wl = pkg.configs.tiny(n=32, w=16, h=16, spp=8, scattevs=3)
    k = wl.kernel()
    rng = np.random.default_rng(3)
    gl = torch.tensor(rng.normal(size=(16, 16, 4)).astype(np.float32), device="cuda")
    ge = torch.tensor(rng.normal(size=(16, 16, 4)).astype(np.float32), device="cuda")
    grad = torch.zeros(wl.scene.leafx[1] * wl.scene.lvoxc3, device="cuda")
    k.adjoint_device(0, 0, 16, 16, 8, 4, 0, gl, ge, grad)
    k.adjoint_device(0, 0, 16, 16, 8, 4, 0, gl, ge, grad)
    torch.cuda.synchronize()
    once = k.adjoint(0, 0, 16, 16, 8, 4, 0, gl.cpu().numpy(), ge.cpu().numpy())
    assert np.allclose(grad.cpu().numpy(), 2 * once, rtol=1e-4, atol=1e-4 * np.abs(once).max())
    k.Destroy()


def _fd_setup(pkg, n=64, w=96, h=96, spp=64, scattevs=8):
    g = pkg.configs.synth_fbm(n, 7)
    sc = pkg.Scene.from_grid(g)
    wl = pkg.configs.Workload("fd", g, sc, w, h, spp, scattevs, 8.0)
    wl.scene = raised_scene(pkg, wl)
    return wl


@pytest.mark.parametrize("direction", ["scale", "random"])
def test_directional_derivative_matches_central_differences(pkg, gpu, direction):
    """<grad, u> against (L(rho + h u) - L(rho - h u)) / 2h with the same seeds."""
    import torch
    wl = _fd_setup(pkg)
    sc = wl.scene
    rho = sc.pvol[1].copy()
    rng = np.random.default_rng(11)
    u = rho.copy() if direction == "scale" else (rho * rng.uniform(0, 1, rho.size)).astype(np.float32)
    gl = np.zeros((wl.h, wl.w, 4), np.float32)
    gl[..., :3] = 1000.0
    gl[..., 3] = 0.01
    ge = np.zeros_like(gl)
    ge[..., :3] = 1.0
    k = wl.kernel()
    n = wl.w * wl.h * 4
    out_l = torch.empty(n, device="cuda")
    out_e = torch.empty(n, device="cuda")
    tgl, tge = torch.tensor(gl, device="cuda").reshape(-1), torch.tensor(ge, device="cuda").reshape(-1)
    h = 0.05

    def loss(pvol, seed):
        sc.update_fog(pvol)
        k.render_device(0, 0, wl.w, wl.h, wl.spp, seed, 0, out_l, out_e)
        torch.cuda.synchronize()
        return float((out_l.double() * tgl.double()).sum() + (out_e.double() * tge.double()).sum())
    fd, ad = [], []
    for s in range(12):
        fd.append((loss(rho + h * u, s) - loss(rho - h * u, s)) / (2 * h))
        sc.update_fog(rho)
        ad.append(float(k.adjoint(0, 0, wl.w, wl.h, wl.spp, s, 0, gl, ge).astype(np.float64) @ u.astype(np.float64)))
    k.Destroy()
    fd, ad = np.array(fd), np.array(ad)
    se = math.sqrt(fd.var(ddof=1) / len(fd) + ad.var(ddof=1) / len(ad))
    z = (ad.mean() - fd.mean()) / se
    assert abs(z) < 4.0, (ad.mean(), fd.mean(), se)
    assert abs(ad.mean() / fd.mean() - 1) < 0.05, (ad.mean(), fd.mean())


def test_autograd_function(pkg, gpu):
    """torch.autograd wrapper: backward equals the adjoint kernel for the
    same keys, and a gradient step lowers an L2 image loss."""
    import torch
    from droplet_render_amd import autograd as AG
    wl = pkg.configs.tiny(n=32, w=24, h=24, spp=32, scattevs=3)
    wl.scene = raised_scene(pkg, wl, 2.0)
    k = wl.kernel()
    rho0 = torch.tensor(wl.scene.pvol[1], device="cuda")
    target_l, target_e = AG.render(k, rho0, samples=128, seed=99)
    target_l, target_e = (target_l / 128).detach(), (target_e / 128).detach()
    rho = (rho0 * 0.6).clone().requires_grad_(True)
    il, ie = AG.render(k, rho, samples=wl.spp, seed=1)
    loss = ((ie[..., :3] / wl.spp - target_e[..., :3]) ** 2).sum()
    loss.backward()
    g = rho.grad.detach().cpu().numpy()
    ge = (2 * (ie[..., :3] / wl.spp - target_e[..., :3]) / wl.spp).detach()
    adj_e = torch.zeros(wl.h, wl.w, 4, device="cuda")
    adj_e[..., :3] = ge
    ref = k.adjoint(0, 0, wl.w, wl.h, wl.spp, 1, 0, np.zeros((wl.h, wl.w, 4), np.float32), adj_e.cpu().numpy())
    # float atomics add in a different order on each run
    assert np.linalg.norm(g - ref) <= 1e-4 * np.linalg.norm(ref), np.linalg.norm(g - ref) / np.linalg.norm(ref)
    # a small step against the gradient reduces the loss (same seed)
    with torch.no_grad():
        step = rho - 0.2 * rho.grad / rho.grad.abs().max()
        step.clamp_(min=0.0)
        il2, ie2 = AG.render(k, step.contiguous(), samples=wl.spp, seed=1)
        loss2 = ((ie2[..., :3] / wl.spp - target_e[..., :3]) ** 2).sum()
    assert float(loss2) < float(loss.detach())
    k.Destroy()


@pytest.mark.parametrize("scattevs", [3, 40])
def test_record_adjoint_matches_replay(pkg, gpu, scattevs):
    """Forward with the path record writes the same sums as the plain forward,
    and the adjoint from the record equals the two-pass replay (same per-path
    terms; only the order of the float atomics differs). scattevs 40 uses the
    high word of the shadow-escape bits."""
    import torch
    from droplet_render_amd import _lib as L
    wl = pkg.configs.tiny(n=32, w=40, h=36, spp=12, scattevs=scattevs)
    wl.scene = raised_scene(pkg, wl)
    k = wl.kernel()
    x0, y0, tx, ty = 16, 10, 30, 28        # a sub-tile clipped by the image edge to 24 x 26
    n = tx * ty * 4
    ref_l, ref_e = torch.empty(n, device="cuda"), torch.empty(n, device="cuda")
    out_l, out_e = torch.empty(n, device="cuda"), torch.empty(n, device="cuda")
    nb = k.record_bytes(x0, y0, tx, ty, wl.spp)
    assert nb == 24 * 26 * wl.spp * 40
    rec = torch.empty(nb // 4, dtype=torch.float32, device="cuda")
    k.render_device(x0, y0, tx, ty, wl.spp, 8, 3, ref_l, ref_e)
    k.render_record_device(x0, y0, tx, ty, wl.spp, 8, 3, out_l, out_e, rec)
    torch.cuda.synchronize()
    assert torch.equal(out_l, ref_l) and torch.equal(out_e, ref_e)
    rng = np.random.default_rng(4)
    gl = torch.tensor(rng.normal(size=n).astype(np.float32) * 100.0, device="cuda")
    ge = torch.tensor(rng.normal(size=n).astype(np.float32), device="cuda")
    g_rep = torch.zeros(wl.scene.leafx[1] * wl.scene.lvoxc3, device="cuda")
    g_rec = torch.zeros_like(g_rep)
    k.adjoint_device(x0, y0, tx, ty, wl.spp, 8, 3, gl, ge, g_rep)
    k.adjoint_record_device(x0, y0, tx, ty, wl.spp, 8, 3, gl, ge, g_rec, rec)
    torch.cuda.synchronize()
    a, b = g_rep.double().cpu().numpy(), g_rec.double().cpu().numpy()
    assert np.isfinite(b).all() and np.linalg.norm(a) > 0
    assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(a), np.linalg.norm(a - b) / np.linalg.norm(a)
    assert np.array_equal(a == 0, b == 0)
    # A record shorter than the call needs is rejected before any launch.
    with pytest.raises(L.DrkError):
        k.adjoint_record_device(x0, y0, tx, ty, wl.spp + 1, 8, 3, gl, ge, g_rec, rec)
    with pytest.raises(L.DrkError):
        k.render_record_device(x0, y0, tx, ty, wl.spp + 1, 8, 3, out_l, out_e, rec)
    k.Destroy()


def test_event_binning_matches_direct_atomics(pkg, gpu):
    """The binned gradient events (sorted by brick, summed in LDS) give the
    same gradient as per-corner float atomics, also when the event buffer
    overflows part way (the rest falls back to atomics)."""
    import torch
    wl = pkg.configs.tiny(n=32, w=48, h=40, spp=16, scattevs=6)
    wl.scene = raised_scene(pkg, wl)
    k = wl.kernel()
    n = wl.w * wl.h * 4
    rng = np.random.default_rng(9)
    gl = torch.tensor(rng.normal(size=n).astype(np.float32) * 100.0, device="cuda")
    ge = torch.tensor(rng.normal(size=n).astype(np.float32), device="cuda")
    gn = wl.scene.leafx[1] * wl.scene.lvoxc3
    out = {}
    for name, budget in [("atomics", 0), ("overflow", 48 * 4096), ("events", None)]:
        k.set_event_budget(budget)
        g = torch.zeros(gn, device="cuda")
        k.adjoint_device(0, 0, wl.w, wl.h, wl.spp, 12, 0, gl, ge, g)
        k.adjoint_device(0, 0, wl.w, wl.h, wl.spp, 13, 0, gl, ge, g)   # accumulates a second call
        torch.cuda.synchronize()
        out[name] = g.double().cpu().numpy()
    k.Destroy()
    ref = out["atomics"]
    assert np.isfinite(ref).all() and np.linalg.norm(ref) > 0
    for name in ("overflow", "events"):
        d = np.linalg.norm(out[name] - ref) / np.linalg.norm(ref)
        assert d <= 1e-5, (name, d)
        assert np.array_equal(out[name] == 0, ref == 0), name


def test_event_binning_with_thousands_of_bins(pkg, gpu):
    """A scene of 32768 bricks (every leaf of a 264^3 lattice occupied) sorts
    its events into 4096 bins: the scatter's bin scan spans several bins per
    thread and its LDS passes 64 KB. Binned events still give the gradient of
    per-corner float atomics."""
    import torch
    g = np.zeros((264, 264, 264), np.float32)
    g[::3, ::3, ::3] = 0.5
    sc = pkg.scene.Scene.from_grid(g)
    assert sc.leafx[1] == 32768
    wl = pkg.configs.Workload("lattice264", g, sc, 48, 40, 8, 4, 8.0)
    k = wl.kernel()
    n = wl.w * wl.h * 4
    rng = np.random.default_rng(4)
    gl = torch.tensor(rng.normal(size=n).astype(np.float32) * 100.0, device="cuda")
    ge = torch.tensor(rng.normal(size=n).astype(np.float32), device="cuda")
    gn = sc.leafx[1] * sc.lvoxc3
    out = {}
    for name, budget in [("atomics", 0), ("events", None)]:
        k.set_event_budget(budget)
        gr = torch.zeros(gn, device="cuda")
        k.adjoint_device(0, 0, wl.w, wl.h, wl.spp, 21, 0, gl, ge, gr)
        torch.cuda.synchronize()
        out[name] = gr.double().cpu().numpy()
    k.Destroy()
    ref = out["atomics"]
    assert np.isfinite(ref).all() and np.count_nonzero(ref) > 10000
    d = np.linalg.norm(out["events"] - ref) / np.linalg.norm(ref)
    assert d <= 1e-5, d
    # Zeros agree except for entries whose splats are below the fixed-point
    # quantum of their chunk (2^-44 of the chunk's largest weight): a corner
    # weight (1 - nl)(1 - nl') of ~1e-14 rounds to zero there.
    tiny = np.abs(ref) <= 2.0 ** -40 * np.abs(ref).max()
    assert np.all(((out["events"] == 0) == (ref == 0)) | tiny)


def test_event_binning_keeps_non_finite_weights(pkg, gpu):
    """A NaN in the image gradient makes the weights of that pixel's events
    NaN: the binned path (whose fixed-point sums cannot hold one) must carry
    them into the same gradient entries as the float atomics do, and leave
    every other entry as the atomics leave it."""
    import torch
    wl = pkg.configs.tiny(n=32, w=32, h=32, spp=8, scattevs=4)
    wl.scene = raised_scene(pkg, wl)
    k = wl.kernel()
    n = wl.w * wl.h * 4
    rng = np.random.default_rng(4)
    gl_h = rng.normal(size=n).astype(np.float32)
    gl_h[(16 * wl.w + 16) * 4] = np.nan   # red channel of the centre pixel
    gl = torch.tensor(gl_h, device="cuda")
    ge = torch.tensor(rng.normal(size=n).astype(np.float32), device="cuda")
    gn = wl.scene.leafx[1] * wl.scene.lvoxc3
    out = {}
    for name, budget in [("atomics", 0), ("events", None)]:
        k.set_event_budget(budget)
        g = torch.zeros(gn, device="cuda")
        k.adjoint_device(0, 0, wl.w, wl.h, wl.spp, 21, 0, gl, ge, g)
        torch.cuda.synchronize()
        out[name] = g.double().cpu().numpy()
    k.Destroy()
    a, e = out["atomics"], out["events"]
    bad = np.isnan(a)
    assert bad.any() and not bad.all()
    assert np.array_equal(np.isnan(e), bad)
    fin = ~bad
    assert np.linalg.norm(e[fin] - a[fin]) <= 1e-5 * np.linalg.norm(a[fin])

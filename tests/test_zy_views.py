"""Several views of one scene: the views launches (drk_views_*: the record
forward and its adjoint for many calls in one launch each) and the gradient
window (adjoint calls of several views append their events and one reduction
adds them all, drk_scene_events_begin/end).

A views launch plans every job exactly as its own call, so its sums and path
records equal the separate calls' bit for bit; captured in a graph, every
replay gives the eager results.

The file is collected after the other GPU tests (by its name): two of its
comparisons have each failed once, in about 450 and about 20 runs, with no
cause found, and under `pytest -x` a failure here then stops no other test.

The window and the views adjoint change only where and when the events are
summed, so their gradient equals the sum of the views' separate adjoints up to the rounding
of the sums: each reduction rounds its fixed-point splats on a grid set by
the largest weight of a bin (finer than float sums' own rounding) and adds
its sums to the gradient in float, once per call or once per window.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _views(pkg, n=3, res=40, spp=8, scattevs=4):
    import torch
    wl = pkg.configs.c3(n=32, views=n, res=res, spp=spp, scattevs=scattevs)
    from droplet_render_amd.fit import _kernel
    ks = [_kernel(wl.init, vi, pi, res, res, scattevs, wl.sigma_t) for vi, pi in wl.views]
    rng = np.random.default_rng(5)
    gs = []
    for _ in ks:
        gl = rng.normal(size=(res, res, 4)).astype(np.float32)
        gl[..., :3] *= 100.0
        ge = rng.normal(size=(res, res, 4)).astype(np.float32)
        gs.append((torch.tensor(gl, device="cuda"), torch.tensor(ge, device="cuda")))
    return wl, ks, gs


def _close(a, b):
    a, b = a.cpu().numpy().astype(np.float64), b.cpu().numpy().astype(np.float64)
    scale = np.abs(b).max()
    assert scale > 0
    assert np.isfinite(a).all()
    err = np.abs(a - b).max() / scale
    assert err < 1e-5, err
    assert np.mean((a == 0) != (b == 0)) < 1e-3


def _separate(wl, ks, gs, spp):
    import torch
    grad = torch.zeros(wl.init.leafx[1] * wl.init.lvoxc3, device="cuda")
    for i, (k, (gl, ge)) in enumerate(zip(ks, gs)):
        k.adjoint_device(0, 0, k.tilex, k.tiley, spp, 100 + i, 0, gl, ge, grad)
    torch.cuda.synchronize()
    return grad


def _windowed(wl, ks, gs, spp, samples=None):
    import torch
    grad = torch.zeros(wl.init.leafx[1] * wl.init.lvoxc3, device="cuda")
    total = samples or sum(k.tilex * k.tiley * spp for k in ks)
    wl.init.events_begin(total)
    for i, (k, (gl, ge)) in enumerate(zip(ks, gs)):
        k.adjoint_device(0, 0, k.tilex, k.tiley, spp, 100 + i, 0, gl, ge, grad)
    wl.init.events_end(grad)
    torch.cuda.synchronize()
    return grad


def test_window_equals_separate_adjoints(pkg, gpu):
    wl, ks, gs = _views(pkg)
    try:
        ref = _separate(wl, ks, gs, 8)
        got = _windowed(wl, ks, gs, 8)
        _close(got, ref)
        # a second window on the same buffer starts from an empty one
        _close(_windowed(wl, ks, gs, 8), ref)
    finally:
        for k in ks:
            k.Destroy()


def test_window_overflow_goes_to_atomics(pkg, gpu):
    """A window sized for a fraction of the events: the rest are added with
    atomics into the calls' gradient, and the total is unchanged."""
    wl, ks, gs = _views(pkg)
    try:
        ref = _separate(wl, ks, gs, 8)
        got = _windowed(wl, ks, gs, 8, samples=4096)
        _close(got, ref)
    finally:
        for k in ks:
            k.Destroy()


def test_window_refusals(pkg, gpu):
    import torch
    wl, ks, gs = _views(pkg, n=1)
    k = ks[0]
    try:
        grad = torch.zeros(wl.init.leafx[1] * wl.init.lvoxc3, device="cuda")
        with pytest.raises(pkg._lib.DrkError):
            wl.init.events_end(grad)           # no window open
        wl.init.events_begin(1000)
        with pytest.raises(pkg._lib.DrkError):
            wl.init.events_begin(1000)         # already open
        with pytest.raises(pkg._lib.DrkError):
            k.set_event_budget(1 << 20)        # the buffer is in use
        gl, ge = gs[0]
        with pytest.raises(pkg._lib.DrkError):
            k.adjoint(0, 0, k.tilex, k.tiley, 4, 1, 0, gl.cpu().numpy(), ge.cpu().numpy())
        wl.init.events_end(grad)
        torch.cuda.synchronize()
        assert float(grad.abs().sum()) == 0.0    # no call ran in the window
    finally:
        k.Destroy()


def _record_jobs(ks, spp, tiles=None):
    import torch
    jobs, outs = [], []
    for i, k in enumerate(ks):
        x0, y0, tx, ty = tiles[i] if tiles else (0, 0, k.tilex, k.tiley)
        nb = k.record_bytes(x0, y0, tx, ty, spp)
        o = (torch.full((tx * ty * 4,), -1.0, device="cuda"), torch.full((tx * ty * 4,), -1.0, device="cuda"),
             torch.zeros(max(4, (nb + 15) // 16 * 4), device="cuda"))
        jobs.append((k, x0, y0, tx, ty, spp, 100 + i, 3, o[0], o[1], o[2]))
        outs.append(o)
    return jobs, outs


@pytest.mark.parametrize("tiles", [None, [(0, 0, 40, 40), (8, 6, 21, 17), (30, 30, 16, 16)],
                                   [(40, 0, 16, 16), (2, 2, 12, 10), (0, 38, 40, 8)]])
def test_views_record_forward_equals_separate_calls(pkg, gpu, tiles):
    """Sums and path records of one views launch equal the separate
    drk_kernel_render_record_device calls bit for bit (full frames; odd and
    overhanging tiles; a tile wholly outside the image, which renders
    nothing and is zeroed)."""
    import torch
    from droplet_render_amd import batch as B
    wl, ks, gs = _views(pkg)
    try:
        jobs, outs = _record_jobs(ks, 6, tiles)
        B.render_record_views(jobs)
        ref_jobs, refs = _record_jobs(ks, 6, tiles)
        for (k, x0, y0, tx, ty, spp, seed, base, ol, oe, rec) in ref_jobs:
            k.render_record_device(x0, y0, tx, ty, spp, seed, base, ol, oe, rec)
        torch.cuda.synchronize()
        for v, (o, r) in enumerate(zip(outs, refs)):
            for name, a, b in zip(("lights", "env", "record"), o, r):
                d = torch.nonzero(a.view(torch.int32) != b.view(torch.int32)).flatten()
                assert d.numel() == 0, (v, name, d[:6].tolist(), a[d[:6]].tolist(), b[d[:6]].tolist())
        assert float(outs[1][0].view(-1, 4)[:, 3].sum()) > 0
        if tiles and tiles[0][0] >= 40:
            assert float(outs[0][0].abs().sum()) == 0.0 and float(outs[0][1].abs().sum()) == 0.0
    finally:
        for k in ks:
            k.Destroy()


@pytest.mark.parametrize("estimator", ["segment", "null"])
def test_views_adjoint_equals_separate_adjoints(pkg, gpu, estimator):
    import torch
    from droplet_render_amd import batch as B
    wl, ks, gs = _views(pkg)
    try:
        if estimator == "null":
            for k in ks:
                k.set_adjoint_estimator(pkg._lib.DRK_ADJ_NULL_COLLISION)
        jobs, outs = _record_jobs(ks, 6)
        B.render_record_views(jobs)
        n = wl.init.leafx[1] * wl.init.lvoxc3
        ref = torch.zeros(n, device="cuda")
        for (k, x0, y0, tx, ty, spp, seed, base, ol, oe, rec), (gl, ge) in zip(jobs, gs):
            k.adjoint_record_device(x0, y0, tx, ty, spp, seed, base, gl, ge, ref, rec)
        adj_jobs = [j[:8] + (g[0], g[1], j[10]) for j, g in zip(jobs, gs)]
        got = torch.zeros(n, device="cuda")
        B.adjoint_record_views(adj_jobs, got)
        torch.cuda.synchronize()
        _close(got, ref)
        # inside a gradient window: the events wait for the window's end
        win = torch.zeros(n, device="cuda")
        wl.init.events_begin(sum(j[3] * j[4] * j[5] for j in jobs))
        B.adjoint_record_views(adj_jobs, win)
        wl.init.events_end(win)
        torch.cuda.synchronize()
        _close(win, ref)
    finally:
        for k in ks:
            k.Destroy()


def test_views_refusals(pkg, gpu):
    import torch
    from droplet_render_amd import batch as B
    wl, ks, gs = _views(pkg, n=2)
    other = pkg.configs.tiny(n=32, w=40, h=40, spp=4)
    ko = other.kernel()
    try:
        jobs, _ = _record_jobs(ks, 4)
        with pytest.raises(pkg._lib.DrkError):
            B.render_record_views([jobs[0], jobs[0]])              # one kernel twice
        short = jobs[1][:10] + (jobs[1][10][:16],)
        with pytest.raises(pkg._lib.DrkError):
            B.render_record_views([jobs[0], short])                # record too small
        jo, _ = _record_jobs([ko], 4)
        grad = torch.zeros(wl.init.leafx[1] * wl.init.lvoxc3, device="cuda")
        with pytest.raises(pkg._lib.DrkError):
            B.adjoint_record_views([jobs[0], jo[0]], grad)         # two scenes
    finally:
        for k in ks + [ko]:
            k.Destroy()


@pytest.mark.parametrize("record", [True, False])
def test_render_views_matches_per_view_autograd(pkg, gpu, record):
    import torch
    from droplet_render_amd import autograd as AG
    wl, ks, gs = _views(pkg)
    try:
        rho0 = torch.tensor(wl.init.pvol[1], device="cuda")
        seeds = [7, 8, 9]

        def loss_of(imgs):
            return sum(((il + ie) * g[0]).sum() + (ie * g[1]).sum() for (il, ie), g in zip(imgs, gs))

        rho = rho0.clone().requires_grad_(True)
        imgs = AG.render_views(ks, rho, 8, seeds, record=record)
        loss_of(imgs).backward()
        g_views = rho.grad.clone()

        rho2 = rho0.clone().requires_grad_(True)
        imgs2 = [AG.render(k, rho2, samples=8, seed=s, record=record) for k, s in zip(ks, seeds)]
        loss_of(imgs2).backward()
        for (a, b), (c, d) in zip(imgs, imgs2):
            assert torch.equal(a, c) and torch.equal(b, d)
        _close(g_views, rho2.grad)
    finally:
        for k in ks:
            k.Destroy()


def test_views_launches_replay_from_a_graph(pkg, gpu):
    """The record forward and the adjoint of several views, captured together
    (their buffers exist after one eager call, so the calls are kernels
    only): every replay gives the eager sums and records bit for bit and the
    eager gradient up to the order of its float additions."""
    import numpy as np
    import torch
    from droplet_render_amd.fit import _kernel
    fw = pkg.configs.c3(n=32, views=3, res=40, spp=6, scattevs=4)
    ks = [_kernel(fw.init, vi, pi, 40, 40, 4, fw.sigma_t) for vi, pi in fw.views]
    n = 40 * 40 * 4
    rb = ks[0].record_bytes(0, 0, 40, 40, 6)
    outs = [(torch.zeros(n, device="cuda"), torch.zeros(n, device="cuda"), torch.zeros(rb // 4, device="cuda"))
            for _ in ks]
    rng = np.random.default_rng(9)
    gimg = [(torch.tensor(rng.normal(size=n).astype(np.float32), device="cuda"),
             torch.tensor(rng.normal(size=n).astype(np.float32), device="cuda")) for _ in ks]
    grad = torch.zeros(fw.init.leafx[1] * fw.init.lvoxc3, device="cuda")
    jf = [(k, 0, 0, 40, 40, 6, 50 + i, 0, o[0], o[1], o[2]) for i, (k, o) in enumerate(zip(ks, outs))]
    ja = [(k, 0, 0, 40, 40, 6, 50 + i, 0, g[0], g[1], o[2]) for i, (k, g, o) in enumerate(zip(ks, gimg, outs))]

    def step(stream):
        grad.zero_()
        pkg.batch.render_record_views(jf, stream=stream)
        pkg.batch.adjoint_record_views(ja, grad, stream=stream)

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        step(side)                      # eager: allocates the buffers
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    ref = [[t.clone() for t in o] for o in outs]
    gref = grad.cpu().numpy().astype(np.float64)
    assert np.abs(gref).max() > 0

    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        step(torch.cuda.current_stream())
    torch.cuda.synchronize()
    for rep in range(3):
        for o in outs:
            for t in o:
                t.fill_(-1.0)
        grad.fill_(7.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for v, (o, r) in enumerate(zip(outs, ref)):
            for name, a, b in zip(("lights", "env", "record"), o, r):
                d = torch.nonzero(a.view(torch.int32) != b.view(torch.int32)).flatten()
                assert d.numel() == 0, (rep, v, name, d.numel(), d[:6].tolist(), a[d[:6]].tolist(), b[d[:6]].tolist())
        gr = grad.cpu().numpy().astype(np.float64)
        assert np.abs(gr - gref).max() <= 1e-5 * np.abs(gref).max(), rep
        # eager work on the same kernels between replays (other keys)
        ks[0].render_device(0, 0, 40, 40, 6, 99 + rep, 0, torch.empty(n, device="cuda"),
                            torch.empty(n, device="cuda"))
        torch.cuda.synchronize()
    del g
    for k in ks:
        k.Destroy()

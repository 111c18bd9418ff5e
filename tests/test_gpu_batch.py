"""Batched persistent launch: many scenes, views and tiles of uneven size in
one grid must give, per job, exactly the sums of a separate launch with the
same RNG keys and chunking; the captured-graph replay must too, every time."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _jobs(pkg, torch):
    wls = [pkg.configs.tiny(n=32, w=32, h=32, spp=6, scattevs=4),
           pkg.configs.tiny(n=40, w=47, h=33, spp=5, scattevs=3, seed=11),   # odd tile: last row/column dropped
           pkg.configs.Workload("spheres48", pkg.configs.synth_spheres(48, 4, spheres=12),
                                None, 64, 40, 9, 5, 12.0),
           pkg.configs.tiny(n=32, w=16, h=24, spp=3, scattevs=6, seed=5)]
    wls[2].scene = pkg.Scene.from_grid(wls[2].grid)
    kernels = [wl.kernel() for wl in wls]
    for k in kernels:
        k.set_tuning(chunk=2)          # identical chunking in both launch forms
    jobs = []
    for i, (wl, k) in enumerate(zip(wls, kernels)):
        n = wl.w * wl.h * 4
        jobs.append((k, 0, 0, wl.w, wl.h, wl.spp, 17 + i, 3 * i, torch.empty(n, device="cuda"),
                     torch.empty(n, device="cuda")))
    # a second job on the first kernel: a sub-tile with its own seed
    k0 = kernels[0]
    jobs.append((k0, 8, 4, 20, 18, 4, 99, 0, torch.empty(32 * 32 * 4, device="cuda"),
                 torch.empty(32 * 32 * 4, device="cuda")))
    return wls, kernels, jobs


def _singles(jobs, torch):
    out = []
    for (k, x0, y0, tx, ty, spp, seed, base, ol, oe) in jobs:
        rl, re = torch.empty_like(ol), torch.empty_like(oe)
        k.render_device(x0, y0, tx, ty, spp, seed, base, rl, re)
        out.append((rl, re))
    torch.cuda.synchronize()
    return out


def test_batch_matches_separate_launches(pkg, gpu):
    torch = gpu
    wls, kernels, jobs = _jobs(pkg, torch)
    ref = _singles(jobs, torch)
    b = pkg.batch.Batch(jobs)
    b.launch()
    torch.cuda.synchronize()
    for (k, x0, y0, tx, ty, *_r), (rl, re), j in zip(jobs, ref, jobs):
        n = tx * ty * 4
        assert torch.equal(j[8][:n], rl[:n]) and torch.equal(j[9][:n], re[:n])
        assert float(rl[:n].abs().sum()) > 0
    b.destroy()
    for k in kernels:
        k.Destroy()


def test_batch_graph_replays_are_identical(pkg, gpu):
    torch = gpu
    wls, kernels, jobs = _jobs(pkg, torch)
    ref = _singles(jobs, torch)
    b = pkg.batch.Batch(jobs)
    for rep in range(3):
        for j in jobs:
            j[8].fill_(-1.0)
            j[9].fill_(-1.0)
        b.launch_graph()
        torch.cuda.synchronize()
        junk = torch.randn(1 << 16, device="cuda").sum()   # eager work between replays
        torch.cuda.synchronize()
        for j, (rl, re) in zip(jobs, ref):
            n = j[3] * j[4] * 4
            assert torch.equal(j[8][:n], rl[:n]) and torch.equal(j[9][:n], re[:n]), (rep, j[3], j[4])
    assert bool(torch.isfinite(junk))
    b.destroy()
    for k in kernels:
        k.Destroy()


def test_autotune_keeps_results(pkg, gpu):
    torch = gpu
    wls, kernels, jobs = _jobs(pkg, torch)
    ref = _singles(jobs, torch)
    b = pkg.batch.Batch(jobs)
    report = b.autotune(reps=1)
    lines = report.strip().splitlines()
    assert lines[0].startswith("grid") and lines[-1].startswith("chosen") and len(lines) == 12
    b.launch()
    torch.cuda.synchronize()
    for j, (rl, re) in zip(jobs, ref):
        n = j[3] * j[4] * 4
        assert torch.equal(j[8][:n], rl[:n]) and torch.equal(j[9][:n], re[:n])
    b.destroy()
    for k in kernels:
        k.Destroy()


def test_batch_rejects_bad_jobs(pkg, gpu):
    torch = gpu
    from droplet_render_amd import _lib as L
    wl = pkg.configs.tiny(n=32, w=16, h=16, spp=2, scattevs=2)
    k = wl.kernel()
    big = torch.empty(64 * 64 * 4, device="cuda")
    with pytest.raises(L.DrkError):
        pkg.batch.Batch([(k, 0, 0, 32, 32, 2, 1, 0, big, big)])     # tile larger than the kernel's
    with pytest.raises(L.DrkError):
        pkg.batch.Batch([(k, 0, 0, 16, 16, 0, 1, 0, big, big)])     # no samples
    k.Destroy()


@pytest.mark.parametrize("chunk", [1, 3])
def test_ray_end_services_change_no_bits(pkg, gpu, chunk):
    """The single-launch kernel runs the scattering phase and the escape
    radiance as workgroup services (an escaping lane's radiance is added to
    its unit's sums in memory by another wave); the batched kernel and the
    allocation-free entry point run them in place. Sums and sums of squares
    agree bit for bit, with single-sample units (every escape after
    scattering deferred) and with chunks of 3 (only escapes that end a unit
    whose sums are still zero are deferred); so do the sums of the forward
    that keeps the path record."""
    torch = gpu
    wl = pkg.configs.tiny(n=40, w=48, h=40, spp=12, scattevs=6, seed=7)
    k = wl.kernel()
    k.set_tuning(chunk=chunk)
    n = wl.w * wl.h * 4

    def bufs():
        return [torch.full((n,), -1.0, device="cuda") for _ in range(4)]
    single, noalloc, batched = bufs(), bufs(), bufs()
    k.render_device(0, 0, wl.w, wl.h, wl.spp, 123, 0, *single)
    k.render_device_noalloc(0, 0, wl.w, wl.h, wl.spp, 123, 0, *noalloc)
    b = pkg.batch.Batch([(k, 0, 0, wl.w, wl.h, wl.spp, 123, 0, *batched)])
    b.launch()
    rec = torch.empty(k.record_bytes(0, 0, wl.w, wl.h, wl.spp) // 4, device="cuda")
    rl, re = torch.empty(n, device="cuda"), torch.empty(n, device="cuda")
    k.render_record_device(0, 0, wl.w, wl.h, wl.spp, 123, 0, rl, re, rec)
    torch.cuda.synchronize()
    for i, (a, c, d) in enumerate(zip(single, batched, noalloc)):
        assert torch.equal(a, c), ("batched", i)
        assert torch.equal(a, d), ("noalloc", i)
    assert torch.equal(rl, single[0]) and torch.equal(re, single[1])
    a = single[0].view(-1, 4)
    assert float(a[:, 3].sum()) > 0 and float(single[1].view(-1, 4)[:, :3].sum()) > 0
    b.destroy()
    k.Destroy()

"""The batched (C2) and large (C4) workloads at their configured sizes.

C2: the 32-scene batch (grids 32^3..488^3, occupancy 2-30%, images 128^2 to
1024^2, 16-64 spp, 16 events) in one persistent launch and as a replayed
hipGraph: every job's sums equal its separate launch bit for bit (same keys,
same chunking), and the sums are well formed.
C4: the 512^3 sparse cloud, 8 views at 1920x1080, 16 spp, 64 events:
full-size properties of the batched forward, forward with path records and
the adjoint into one FP32 gradient; one tile around the cloud against the
oracle with the same keys.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _well_formed(ol, oe, w, h, spp):
    a = ol.view(h, w, 4).cpu().numpy()
    b = oe.view(h, w, 4).cpu().numpy()
    assert np.isfinite(a).all() and np.isfinite(b).all()
    al = a[..., 3]
    assert np.array_equal(al, b[..., 3])
    assert np.array_equal(al, np.round(al)) and al.min() >= 0 and al.max() <= spp
    assert (a[..., :3] >= 0).all() and (b[..., :3] >= 0).all()
    return a, b


def test_c2_batch_jobs_equal_separate_launches(pkg, gpu):
    torch = gpu
    wls = pkg.configs.c2()
    assert len(wls) == 32
    ks = [wl.kernel() for wl in wls]
    for k in ks:
        k.set_tuning(chunk=4)        # the same fixed chunking in both launch forms
    jobs, ref = [], []
    for i, (wl, k) in enumerate(zip(wls, ks)):
        n = wl.w * wl.h * 4
        jobs.append((k, 0, 0, wl.w, wl.h, wl.spp, 700 + i, 0, torch.empty(n, device="cuda"),
                     torch.empty(n, device="cuda")))
        rl, re = torch.empty(n, device="cuda"), torch.empty(n, device="cuda")
        k.render_device(0, 0, wl.w, wl.h, wl.spp, 700 + i, 0, rl, re)
        ref.append((rl, re))
    torch.cuda.synchronize()
    b = pkg.batch.Batch(jobs)
    for form in ("launch", "graph", "graph"):
        for j in jobs:
            j[8].fill_(-1.0)
            j[9].fill_(-1.0)
        (b.launch if form == "launch" else b.launch_graph)()
        torch.cuda.synchronize()
        for i, (j, (rl, re)) in enumerate(zip(jobs, ref)):
            assert torch.equal(j[8], rl) and torch.equal(j[9], re), (form, i)
    lit = 0
    for wl, (rl, re) in zip(wls, ref):
        a, _ = _well_formed(rl, re, wl.w, wl.h, wl.spp)
        lit += a[..., 3].sum() > 0
    assert lit == 32
    b.destroy()
    for k in ks:
        k.Destroy()


def test_c4_full_size(pkg, oracle, gpu):
    torch = gpu
    from droplet_render_amd.fit import _kernel
    wl, views = pkg.configs.c4()
    assert (wl.w, wl.h, wl.spp, wl.scattevs, len(views)) == (1920, 1080, 16, 64, 8)
    ks = [_kernel(wl.scene, vi, pi, wl.w, wl.h, wl.scattevs, wl.sigma_t) for vi, pi in views]
    n = wl.w * wl.h * 4
    outs = [(torch.empty(n, device="cuda"), torch.empty(n, device="cuda")) for _ in ks]
    jobs = [(k, 0, 0, wl.w, wl.h, wl.spp, 900 + i, 0, o[0], o[1]) for i, (k, o) in enumerate(zip(ks, outs))]
    b = pkg.batch.Batch(jobs)
    b.launch()
    torch.cuda.synchronize()
    imgs = [_well_formed(ol, oe, wl.w, wl.h, wl.spp) for ol, oe in outs]
    cover = [float((a[..., 3] > 0).mean()) for a, _ in imgs]
    assert all(0.0005 < c < 0.5 for c in cover), cover
    # forward with path records, then the adjoint of every view into one gradient
    rb = ks[0].record_bytes(0, 0, wl.w, wl.h, wl.spp)
    rec = torch.empty(rb // 4, device="cuda")
    g = torch.Generator(device="cpu").manual_seed(3)
    adj_l = torch.rand(n, generator=g).cuda()
    adj_e = torch.rand(n, generator=g).cuda()
    grad = torch.zeros(wl.scene.leafx[1] * wl.scene.lvoxc3, device="cuda")
    sep = []
    for i, k in enumerate(ks):
        rl, re = torch.empty(n, device="cuda"), torch.empty(n, device="cuda")
        k.render_record_device(0, 0, wl.w, wl.h, wl.spp, 900 + i, 0, rl, re, rec)
        torch.cuda.synchronize()
        # the record-keeping forward renders the same sums (alpha exactly)
        assert torch.equal(rl.view(-1, 4)[:, 3], outs[i][0].view(-1, 4)[:, 3])
        assert torch.allclose(rl, outs[i][0], rtol=1e-5, atol=1e-6)
        k.adjoint_record_device(0, 0, wl.w, wl.h, wl.spp, 900 + i, 0, adj_l, adj_e, grad, rec)
        sep.append((rl, re))
    torch.cuda.synchronize()
    gr = grad.cpu().numpy()
    assert np.isfinite(gr).all() and np.abs(gr).max() > 0
    # the same 8 record forwards and adjoints as two views launches
    del rec
    recs = [torch.empty(rb // 4, device="cuda") for _ in ks]
    vo = [(torch.empty(n, device="cuda"), torch.empty(n, device="cuda")) for _ in ks]
    pkg.batch.render_record_views([(k, 0, 0, wl.w, wl.h, wl.spp, 900 + i, 0, vo[i][0], vo[i][1], recs[i])
                                   for i, k in enumerate(ks)])
    grad2 = torch.zeros_like(grad)
    pkg.batch.adjoint_record_views([(k, 0, 0, wl.w, wl.h, wl.spp, 900 + i, 0, adj_l, adj_e, recs[i])
                                    for i, k in enumerate(ks)], grad2)
    torch.cuda.synchronize()
    for (a, b_), (c, d) in zip(vo, sep):
        assert torch.equal(a, c) and torch.equal(b_, d)
    g2 = grad2.cpu().numpy()
    assert np.abs(g2.astype(np.float64) - gr).max() <= 1e-5 * np.abs(gr).max()
    del recs
    fp = wl.scene.footprint() + sum(k.footprint() for k in ks)
    print("c4 coverage", np.round(cover, 4), "footprint GB %.2f" % (fp / 1e9))
    # one 48x48 tile around the cloud of view 0 against the oracle (same keys)
    a0 = imgs[0][0][..., 3]
    ys, xs = np.nonzero(a0)
    cy, cx = int(np.median(ys)), int(np.median(xs))
    x0, y0 = max(0, min(wl.w - 48, cx - 24)) & ~1, max(0, min(wl.h - 48, cy - 24)) & ~1
    wl0 = pkg.configs.Workload(wl.name, wl.grid, wl.scene, wl.w, wl.h, wl.spp, wl.scattevs, wl.sigma_t,
                               viewi=views[0][0], proji=views[0][1])
    c = oracle.render(wl0, seed=900, x0=x0, y0=y0, tilex=48, tiley=48)
    t = [imgs[0][0][y0:y0 + 48, x0:x0 + 48], imgs[0][1][y0:y0 + 48, x0:x0 + 48]]
    assert c[0][..., 3].sum() > 0
    for bi in range(2):
        assert np.array_equal(t[bi][..., 3], c[bi][..., 3])
        rel = np.abs(t[bi][..., :3].astype(np.float64) - c[bi][..., :3]) / np.maximum(np.abs(c[bi][..., :3]), 1e-6)
        assert np.mean(rel.max(axis=2) < 1e-4) >= 0.99, np.mean(rel.max(axis=2) < 1e-4)
    b.destroy()
    for k in ks:
        k.Destroy()

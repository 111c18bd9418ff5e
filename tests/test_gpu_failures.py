"""Failure handling through the C ABI, forced from the host: the watchdog's
round limit (DRK_ELIMIT, every kernel still ends normally) and device
allocation failures injected with the allocation gate (DRK_ENOMEM, nothing
leaked, the library stays usable; the adjoint's event buffer falls back to
atomics)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _tiny(pkg):
    return pkg.configs.tiny(w=32, h=24, spp=8, scattevs=4)


def test_round_limit_returns_elimit(pkg, gpu):
    wl = _tiny(pkg)
    k = wl.kernel()
    k.set_seed(5, 0)
    k.Render(0, 0, wl.w, wl.h, wl.spp)
    ref = k.phb
    k.set_round_limit(3)
    k.set_seed(5, 0)
    with pytest.raises(pkg._lib.DrkError) as e:
        k.Render(0, 0, wl.w, wl.h, wl.spp)
    assert e.value.code == pkg._lib.DRK_ELIMIT
    assert k.limit_hit()
    k.set_round_limit(1)     # untraced Shadow lanes finish a path per round
    with pytest.raises(pkg._lib.DrkError) as e:
        k.Shadow(0, 0, wl.w, wl.h, wl.spp)
    assert e.value.code == pkg._lib.DRK_ELIMIT
    # a limit no wave reaches changes nothing
    k.set_round_limit(1 << 30)
    k.set_seed(5, 0)
    k.Render(0, 0, wl.w, wl.h, wl.spp)
    assert not k.limit_hit()
    for a, b in zip(k.phb, ref):
        assert np.array_equal(a, b)
    k.set_round_limit(0)
    k.set_seed(5, 0)
    k.Render(0, 0, wl.w, wl.h, wl.spp)
    for a, b in zip(k.phb, ref):
        assert np.array_equal(a, b)
    k.Destroy()


def test_round_limit_device_form_and_adjoint(pkg, gpu):
    import torch
    wl = _tiny(pkg)
    k = wl.kernel()
    n = wl.w * wl.h * 4
    bufs = [torch.zeros(n, device="cuda") for _ in range(2)]
    k.set_round_limit(2)
    k.render_device(0, 0, wl.w, wl.h, wl.spp, 7, 0, bufs[0], bufs[1])
    torch.cuda.synchronize()
    assert k.limit_hit()
    adj = np.ones((wl.h, wl.w, 4), np.float32)
    with pytest.raises(pkg._lib.DrkError) as e:
        k.adjoint(0, 0, wl.w, wl.h, wl.spp, 7, 0, adj, adj)
    assert e.value.code == pkg._lib.DRK_ELIMIT
    k.set_round_limit(0)
    k.render_device(0, 0, wl.w, wl.h, wl.spp, 7, 0, bufs[0], bufs[1])
    torch.cuda.synchronize()
    assert not k.limit_hit()
    k.Destroy()


def _workflow(pkg, wl):
    """Scene upload, kernel creation, a chunked render, an adjoint."""
    sc = pkg.Scene(wl.scene.ob, wl.scene.pvol[1], wl.scene.lvoxc)
    k = pkg.RenderKernel()
    try:
        k.Initialize(sc, wl.viewi, wl.proji, wl.phase, wl.env, None, wl.scattevs, wl.msigmas, wl.msigmaa, wl.w,
                     wl.h, wl.w, wl.h, 0, wl.lights)
        k.set_seed(3, 0)
        k.Render(0, 0, wl.w, wl.h, wl.spp)
        img = k.phb
        adj = np.full((wl.h, wl.w, 4), 0.5, np.float32)
        g = k.adjoint(0, 0, wl.w, wl.h, wl.spp, 3, 0, adj, adj)
        return img, g
    finally:
        k.Destroy()
        sc.destroy()


def test_allocation_failures_are_clean(pkg, gpu):
    """Refuses each device allocation of the workflow in turn: every call
    either fails with DRK_ENOMEM or, for the adjoint's event buffer, falls
    back to atomics with the same gradient; no device memory is left behind
    and the next unrestricted run gives the same results."""
    import torch
    L = pkg._lib
    wl = _tiny(pkg)
    sizes = []
    L.set_alloc_gate(lambda nb: sizes.append(nb) and False)
    try:
        img0, g0 = _workflow(pkg, wl)
    finally:
        L.set_alloc_gate(None)
    assert len(sizes) >= 8, sizes
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    fell_back = 0
    for i in range(len(sizes)):
        seen = [0]

        def refuse(nb, i=i, seen=seen):
            seen[0] += 1
            return seen[0] - 1 == i
        L.set_alloc_gate(refuse)
        try:
            img, g = _workflow(pkg, wl)
            # the only refusal that does not fail the call: the event buffer
            fell_back += 1
            for a, b in zip(img, img0):
                assert np.array_equal(a, b)
            assert np.allclose(g, g0, rtol=1e-3, atol=1e-5 * np.abs(g0).max())
        except L.DrkError as e:
            assert e.code == L.DRK_ENOMEM, (i, e)
        finally:
            L.set_alloc_gate(None)
    assert fell_back >= 1
    torch.cuda.synchronize()
    # nothing left allocated (the runtime hands memory out in pages of a few MiB)
    assert torch.cuda.mem_get_info()[0] >= free0 - (16 << 20)
    img, g = _workflow(pkg, wl)
    for a, b in zip(img, img0):
        assert np.array_equal(a, b)

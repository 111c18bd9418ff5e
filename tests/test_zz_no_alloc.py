"""The allocation-free render entry point (drk_kernel_render_device_noalloc)
against drk_kernel_render_device: same sums (bit-identical: same chunking,
same kernels) on two calls in a row, the second with a smaller tile; no
device allocation during the calls (the allocation gate refuses them all);
oversized calls are refused with a clear error."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bufs(torch, n, k=4):
    return [torch.full((n,), -7.0, device="cuda") for _ in range(k)]


def test_noalloc_matches_render_device(pkg, gpu):
    import torch
    L = pkg._lib
    wl = pkg.configs.tiny(w=64, h=48, spp=24, scattevs=5)
    k = wl.kernel()
    calls = [(0, 0, 64, 48, 24, 11, 0), (8, 6, 31, 19, 7, 12, 3)]   # the second smaller (and odd)
    n = 64 * 48 * 4
    for sq in (False, True):
        for (x0, y0, tw, th, spp, seed, base) in calls:
            ref = _bufs(torch, n)
            k.render_device(x0, y0, tw, th, spp, seed, base, ref[0], ref[1], *(ref[2:] if sq else []))
            torch.cuda.synchronize()
            got = _bufs(torch, n)
            refused = []
            L.set_alloc_gate(lambda nb: refused.append(nb) or True)   # any device allocation fails
            try:
                k.render_device_noalloc(x0, y0, tw, th, spp, seed, base, got[0], got[1], *(got[2:] if sq else []))
                torch.cuda.synchronize()
            finally:
                L.set_alloc_gate(None)
            assert refused == []
            m = tw * th * 4
            for b in range(4 if sq else 2):
                a = ref[b][:m].cpu().numpy()
                g = got[b][:m].cpu().numpy()
                assert np.array_equal(a, g), (sq, x0, b)
            assert got[0][:m].cpu().numpy().reshape(th, tw, 4)[..., 3].sum() > 0
    k.Destroy()


def test_noalloc_refuses_oversized_calls(pkg, gpu):
    import torch
    L = pkg._lib
    wl = pkg.configs.tiny(w=2048, h=1024, spp=64, scattevs=1)
    k = wl.kernel()
    bufs = _bufs(torch, 2048 * 1024 * 4, 2)
    with pytest.raises(L.DrkError) as e:      # 2M pixels x 64 chunks x 32 B: beyond the fixed scratch
        k.render_device_noalloc(0, 0, 2048, 1024, 64, 1, 0, bufs[0], bufs[1])
    assert e.value.code == L.DRK_EINVAL and "scratch" in str(e.value)
    k.set_tuning(chunk=64)                    # one chunk per pixel: no partial sums
    k.render_device_noalloc(0, 0, 256, 128, 4, 1, 0, bufs[0], bufs[1])
    torch.cuda.synchronize()
    k.Destroy()

"""The frame driver's tile layout and order (CPU), and its passes against
direct renders of each tile (GPU)."""
import numpy as np
import pytest


def test_tile_grid_covers_the_frame_once(pkg):
    from droplet_render_amd import driver
    for (w, h, tw, th) in [(100, 70, 32, 32), (64, 64, 64, 64), (130, 50, 48, 20)]:
        cover = np.zeros((h, w), int)
        for (x, y, a, b) in driver.tile_grid(w, h, tw, th):
            assert a > 0 and b > 0 and x >= 0 and y >= 0 and x + a <= w and y + b <= h
            cover[y:y + b, x:x + a] += 1
        assert (cover == 1).all()


def test_tile_order_starts_at_the_centre(pkg):
    from droplet_render_amd import driver
    t = driver.tile_grid(96, 96, 32, 32)
    o = driver.tile_order(t, 96, 96, 32, 32)
    assert o[0] == (32, 32, 32, 32) and sorted(o) == sorted(t)


@pytest.mark.gpu
def test_frame_passes_match_tile_renders(pkg, gpu):
    from droplet_render_amd import driver
    torch = gpu
    wl = pkg.configs.tiny(n=32, w=96, h=64, spp=10, scattevs=3)
    k = wl.kernel(tilex=32, tiley=32)
    k.set_seed(21, 0)
    frame = driver.render_frame(k, wl.w, wl.h, 32, 32, 10, 4)
    # Direct renders of each tile with the same sample indices: the kernel's
    # sample counter advances by 10 per tile, in the driver's tile order.
    order = driver.tile_order(driver.tile_grid(wl.w, wl.h, 32, 32), wl.w, wl.h, 32, 32)
    ref = np.zeros((wl.h, wl.w, 4))
    for i, (x0, y0, tw, th) in enumerate(order):
        n = tw * th * 4
        ol, oe = torch.empty(n, device="cuda"), torch.empty(n, device="cuda")
        k.render_device(x0, y0, tw, th, 10, 21, 10 * i, ol, oe)
        torch.cuda.synchronize()
        s = (ol + oe).view(th, tw, 4).double().cpu().numpy()
        ref[y0:y0 + th, x0:x0 + tw] = s * np.array([0.1, 0.1, 0.1, 0.05])
    k.Destroy()
    assert np.allclose(frame["COMBINED"], ref, rtol=1e-5, atol=1e-7)
    assert np.allclose(frame["TRANSMISSION_DIRECT"] + frame["TRANSMISSION_INDIRECT"], frame["COMBINED"][..., :3],
                       rtol=1e-9, atol=1e-12)
    assert frame["COMBINED"][..., 3].max() <= 1.0 and frame["COMBINED"][..., :3].max() > 0

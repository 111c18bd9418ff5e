"""Render launches captured in a HIP graph and replayed.

A render launch is kernels only -- the launch-block setup (counters and
parameters), the camera-ray pass, the path tracer and the chunk reduction --
with no allocation once a handle's buffers exist (and none ever for the
allocation-free entry), so a caller can capture it on its stream and replay
it. Every replay must give the eager call's sums bit for bit, with eager work
on the same handle and synchronizations between the replays.
"""
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("entry", ["render_device", "render_device_noalloc"])
def test_graph_replays_equal_the_eager_call(pkg, gpu, entry):
    import torch
    wl = pkg.configs.tiny(w=48, h=40, spp=12, scattevs=4)
    k = wl.kernel()
    fn = getattr(k, entry)
    n = wl.w * wl.h * 4
    args = (0, 0, wl.w, wl.h, wl.spp, 5, 3)
    ref = [torch.zeros(n, device="cuda") for _ in range(2)]
    fn(*args, ref[0], ref[1])
    torch.cuda.synchronize()
    assert ref[0].view(wl.h, wl.w, 4)[..., 3].sum().item() > 0

    out = [torch.zeros(n, device="cuda") for _ in range(2)]
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):   # warm-up on a side stream, as before any capture
        fn(*args, out[0], out[1], stream=side)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()

    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        fn(*args, out[0], out[1], stream=torch.cuda.current_stream())
    torch.cuda.synchronize()
    other = [torch.zeros(n, device="cuda") for _ in range(2)]
    for rep in range(3):
        out[0].fill_(-1.0)
        out[1].fill_(-1.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for b in range(2):
            assert torch.equal(out[b], ref[b]), (entry, rep, b)
        # eager work on the same handle between replays (other keys)
        fn(0, 0, wl.w, wl.h, wl.spp, 6 + rep, 0, other[0], other[1])
        torch.cuda.synchronize()
        assert not torch.equal(other[0], ref[0])
    del g
    k.Destroy()


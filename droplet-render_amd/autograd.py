"""torch.autograd wrapper of the render and its adjoint.

`render(kernel, density, ...)` returns the two phb sum images for FOG brick
values `density` (a CUDA tensor in the layout of Scene.pvol[FOG]); its
backward runs the path-replay adjoint kernel with the same RNG keys, so
d(image)/d(density) is differentiated for the very samples the forward drew.
With record=True (the default) the forward keeps each path's radiance and
shadow-escape bits (40 bytes per pixel sample, drk_kernel_render_record_device)
until backward, which then traces every path once instead of twice
(drk_kernel_adjoint_record_device); record=False keeps no per-sample state.
`render_views(kernels, density, samples, seeds)` differentiates the full-frame
renders of several kernels of one scene (views) as one node: one launch for
the forward of all views and one for their adjoints, with one event reduction.
The octree's majorants stay fixed, as the adjoint assumes: keep the densities
at or below them (`Scene.raise_majorants`).
"""
import torch

from . import _lib as L


class _Render(torch.autograd.Function):
    @staticmethod
    def forward(ctx, density, kernel, x0, y0, tilex, tiley, samples, seed, sample_base, record):
        scene = kernel.scene
        assert density.is_cuda and density.dtype == torch.float32 and density.is_contiguous()
        assert density.numel() == scene.leafx[1] * scene.lvoxc3
        stream = torch.cuda.current_stream(density.device)
        scene.update_fog_device(density, stream=stream)   # device-to-device copy into the scene, in stream order
        n = tilex * tiley * 4
        out_l = torch.empty(n, dtype=torch.float32, device=density.device)
        out_e = torch.empty(n, dtype=torch.float32, device=density.device)
        ctx.rec = None
        if record:
            nb = kernel.record_bytes(x0, y0, tilex, tiley, samples)
            ctx.rec = torch.empty((nb + 15) // 16 * 4, dtype=torch.float32, device=density.device)
            kernel.render_record_device(x0, y0, tilex, tiley, samples, seed, sample_base, out_l, out_e, ctx.rec,
                                        stream=stream)
        else:
            kernel.render_device(x0, y0, tilex, tiley, samples, seed, sample_base, out_l, out_e, stream=stream)
        ctx.kernel = kernel
        ctx.args = (x0, y0, tilex, tiley, samples, seed, sample_base)
        ctx.save_for_backward(density)
        return out_l.view(tiley, tilex, 4), out_e.view(tiley, tilex, 4)

    @staticmethod
    def backward(ctx, g_l, g_e):
        (density,) = ctx.saved_tensors
        kernel = ctx.kernel
        x0, y0, tilex, tiley, samples, seed, sample_base = ctx.args
        stream = torch.cuda.current_stream(density.device)
        kernel.scene.update_fog_device(density, stream=stream)   # the adjoint replays these densities
        g_l = torch.zeros(tiley, tilex, 4, device=density.device) if g_l is None else g_l.contiguous().float()
        g_e = torch.zeros(tiley, tilex, 4, device=density.device) if g_e is None else g_e.contiguous().float()
        grad = torch.zeros_like(density)
        if ctx.rec is not None:
            kernel.adjoint_record_device(x0, y0, tilex, tiley, samples, seed, sample_base, g_l, g_e, grad, ctx.rec,
                                         stream=stream)
            ctx.rec = None
        else:
            kernel.adjoint_device(x0, y0, tilex, tiley, samples, seed, sample_base, g_l, g_e, grad, stream=stream)
        return grad, None, None, None, None, None, None, None, None, None


class _RenderViews(torch.autograd.Function):
    """Several full-frame renders of one scene (views) as one node. With the
    path record, the forward renders every view in one launch and the
    backward runs every view's adjoint in one launch with one event
    reduction (drk_views_*); without it, the views render one by one and
    their adjoints share the scene's gradient window (one reduction)."""

    @staticmethod
    def forward(ctx, density, kernels, samples, seeds, record):
        from . import batch as B
        scene = kernels[0].scene
        assert all(k.scene is scene for k in kernels)
        assert density.is_cuda and density.dtype == torch.float32 and density.is_contiguous()
        assert density.numel() == scene.leafx[1] * scene.lvoxc3
        stream = torch.cuda.current_stream(density.device)
        scene.update_fog_device(density, stream=stream)   # ordered after density's producer
        outs, recs, jobs = [], [], []
        for k, seed in zip(kernels, seeds):
            n = k.tilex * k.tiley * 4
            out_l = torch.empty(n, dtype=torch.float32, device=density.device)
            out_e = torch.empty(n, dtype=torch.float32, device=density.device)
            if record:
                nb = k.record_bytes(0, 0, k.tilex, k.tiley, samples)
                rec = torch.empty((nb + 15) // 16 * 4, dtype=torch.float32, device=density.device)
                jobs.append((k, 0, 0, k.tilex, k.tiley, samples, seed, 0, out_l, out_e, rec))
                recs.append(rec)
            else:
                k.render_device(0, 0, k.tilex, k.tiley, samples, seed, 0, out_l, out_e, stream=stream)
            outs += [out_l.view(k.tiley, k.tilex, 4), out_e.view(k.tiley, k.tilex, 4)]
        if record:
            B.render_record_views(jobs, stream=stream)
        ctx.kernels, ctx.samples, ctx.seeds, ctx.recs = kernels, samples, seeds, recs if record else None
        ctx.save_for_backward(density)
        return tuple(outs)

    @staticmethod
    def backward(ctx, *grads):
        from . import batch as B
        (density,) = ctx.saved_tensors
        kernels, samples = ctx.kernels, ctx.samples
        scene = kernels[0].scene
        stream = torch.cuda.current_stream(density.device)
        scene.update_fog_device(density, stream=stream)   # the adjoint replays these densities
        grad = torch.zeros_like(density)
        views = []
        for i, (k, seed) in enumerate(zip(kernels, ctx.seeds)):
            g_l, g_e = grads[2 * i], grads[2 * i + 1]
            if g_l is None and g_e is None:
                continue
            shape = (k.tiley, k.tilex, 4)
            g_l = torch.zeros(shape, device=density.device) if g_l is None else g_l.contiguous().float()
            g_e = torch.zeros(shape, device=density.device) if g_e is None else g_e.contiguous().float()
            views.append((i, k, seed, g_l, g_e))
        if ctx.recs is not None:
            if views:
                jobs = [(k, 0, 0, k.tilex, k.tiley, samples, seed, 0, g_l, g_e, ctx.recs[i])
                        for i, k, seed, g_l, g_e in views]
                B.adjoint_record_views(jobs, grad, stream=stream)
            ctx.recs = None
            return grad, None, None, None, None
        scene.events_begin(sum(k.tilex * k.tiley * samples for _, k, _, _, _ in views) or 1, stream=stream)
        try:
            for i, k, seed, g_l, g_e in views:
                k.adjoint_device(0, 0, k.tilex, k.tiley, samples, seed, 0, g_l, g_e, grad, stream=stream)
        finally:
            scene.events_end(grad, stream=stream)
        return grad, None, None, None, None


def render_views(kernels, density, samples, seeds, record=True):
    """Differentiable full-frame renders of several kernels of one scene:
    [(lights sums, environment sums)] per kernel, each (tiley, tilex, 4).
    The backward adds every view's gradient with one event reduction."""
    outs = _RenderViews.apply(density, tuple(kernels), int(samples), tuple(int(s) for s in seeds), bool(record))
    return [(outs[2 * i], outs[2 * i + 1]) for i in range(len(kernels))]


def render(kernel, density, samples=None, seed=0, sample_base=0, x0=0, y0=0, tilex=None, tiley=None, record=True):
    """Differentiable render: (lights sums, environment sums), each (tiley, tilex, 4)."""
    tilex = tilex or kernel.tilex
    tiley = tiley or kernel.tiley
    samples = samples or 1
    return _Render.apply(density, kernel, int(x0), int(y0), int(tilex), int(tiley), int(samples), int(seed),
                         int(sample_base), bool(record))


__all__ = ["render", "render_views", "L"]

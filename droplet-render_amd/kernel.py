"""RenderKernel: the tile renderer, on the MI355X.

Python mirror of droplet-render's `RenderKernel` (src/kernel.h:11-59):
Initialize / Render / Shadow / Destroy and the two result buffers `phb`
(BUFFER_LIGHTS, BUFFER_ENVIRONMENT), each tilex*tiley RGBA sums. The work runs
in libdroplet_amd.so's gfx950 kernels; this class only marshals arguments.
"""
import ctypes as C

import numpy as np

from . import _lib as L
from .sampler import sampler_desc


class RenderKernel:
    BUFFER_LIGHTS = 0
    BUFFER_ENVIRONMENT = 1
    BUFFER_COUNT = 2

    def __init__(self):
        self._h = None
        self._keep = None
        self.tilew = self.tileh = 0

    def Initialize(self, scene, viewi, proji, ppf, penv, pdepth, scattevs, msigmas, msigmaa, tilex, tiley, w, h,
                   flags=0, lights=()):
        """RenderKernel::Initialize (src/kernel.cpp:497-534). `lights` replaces
        the process-global BaseLight::lights registry; lights[0] is sampled."""
        self.Destroy()
        viewi = np.ascontiguousarray(viewi, dtype=np.float32).reshape(16)
        proji = np.ascontiguousarray(proji, dtype=np.float32).reshape(16)
        desc, keep = sampler_desc(ppf, penv, lights)
        depth = None if pdepth is None else np.ascontiguousarray(pdepth, dtype=np.float32).reshape(-1)
        hnd = C.c_void_p()
        L.check(L.lib().drk_kernel_create(scene.device(), L.fptr(viewi), L.fptr(proji), C.byref(desc), L.fptr(depth),
                                          int(scattevs), float(msigmas), float(msigmaa), int(tilex), int(tiley),
                                          int(w), int(h), int(flags), C.byref(hnd)))
        self._h = hnd
        self._keep = (scene, keep, viewi, proji, depth)
        self.scene, self.w, self.h = scene, int(w), int(h)
        self.tilex, self.tiley = int(tilex), int(tiley)
        self.scattevs, self.msigmas, self.msigmaa = int(scattevs), float(msigmas), float(msigmaa)
        return True

    def set_seed(self, seed, sample_base=0):
        L.check(L.lib().drk_kernel_set_seed(self._h, int(seed), int(sample_base)))

    def set_tuning(self, chunk=0, grid_blocks=0):
        L.check(L.lib().drk_kernel_set_tuning(self._h, int(chunk), int(grid_blocks)))

    def set_round_limit(self, rounds=0):
        """Watchdog: rounds per wave and launch (0: none); past it the
        synchronous calls raise DrkError(DRK_ELIMIT)."""
        L.check(L.lib().drk_kernel_set_round_limit(self._h, int(rounds)))

    def limit_hit(self):
        """Whether the kernel's last launch stopped at the round limit."""
        hit = C.c_int()
        L.check(L.lib().drk_kernel_limit_hit(self._h, C.byref(hit)))
        return bool(hit.value)

    def set_timing(self, enable=True):
        L.check(L.lib().drk_kernel_set_timing(self._h, 1 if enable else 0))

    def timing(self):
        """(total ms, launches) of the path-tracing kernel since the last call."""
        ms, n = C.c_double(), C.c_uint32()
        L.check(L.lib().drk_kernel_timing(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def Render(self, x0, y0, tilex, tiley, samples):
        """RenderKernel::Render (src/kernel.cpp:536-567); blocks until phb holds the sums."""
        L.check(L.lib().drk_kernel_render(self._h, int(x0), int(y0), int(tilex), int(tiley), int(samples)))
        self.tilew, self.tileh = int(tilex), int(tiley)

    def Shadow(self, x0, y0, tilex, tiley, samples):
        """RenderKernel::Shadow (src/kernel.cpp:569-607)."""
        L.check(L.lib().drk_kernel_shadow(self._h, int(x0), int(y0), int(tilex), int(tiley), int(samples)))
        self.tilew, self.tileh = int(tilex), int(tiley)

    def _buf(self, b):
        # Packed with the stride of the last call's tile width, like the
        # reference's phb (src/kernel.cpp:560-564).
        p = L.lib().drk_kernel_phb(self._h, b)
        n = self.tileh * self.tilew * 4
        return np.ctypeslib.as_array(p, shape=(n,)).reshape(self.tileh, self.tilew, 4).copy()

    @property
    def phb(self):
        """[lights, environment] per-pixel RGBA sums of the last Render (copies)."""
        return [self._buf(0), self._buf(1)]

    @property
    def sumsq(self):
        """Per-pixel sums of squares of the per-sample [lights, environment] RGB."""
        return [self._buf(2), self._buf(3)]

    def render_device(self, x0, y0, tilex, tiley, samples, seed, sample_base, out_l, out_e, sq_l=None, sq_e=None,
                      stream=None):
        """Device-buffer render into torch float32 tensors of tilex*tiley*4 floats (asynchronous)."""
        def ptr(t):
            if t is None:
                return None
            assert t.is_cuda and t.dtype.is_floating_point and t.is_contiguous() and t.numel() >= tilex * tiley * 4
            return C.c_void_p(t.data_ptr())
        s = None if stream is None else C.c_void_p(stream.cuda_stream)
        L.check(L.lib().drk_kernel_render_device(self._h, int(x0), int(y0), int(tilex), int(tiley), int(samples),
                                                 int(seed), int(sample_base), ptr(out_l), ptr(out_e), ptr(sq_l),
                                                 ptr(sq_e), s))

    def render_device_noalloc(self, x0, y0, tilex, tiley, samples, seed, sample_base, out_l, out_e, sq_l=None,
                              sq_e=None, stream=None):
        """render_device through the allocation-free entry point (its scratch
        is the library's fixed per-device array; calls on one GPU must not
        overlap)."""
        def ptr(t):
            if t is None:
                return None
            assert t.is_cuda and t.dtype.is_floating_point and t.is_contiguous() and t.numel() >= tilex * tiley * 4
            return C.c_void_p(t.data_ptr())
        s = None if stream is None else C.c_void_p(stream.cuda_stream)
        L.check(L.lib().drk_kernel_render_device_noalloc(self._h, int(x0), int(y0), int(tilex), int(tiley),
                                                         int(samples), int(seed), int(sample_base), ptr(out_l),
                                                         ptr(out_e), ptr(sq_l), ptr(sq_e), s))

    def render_weighted_device(self, x0, y0, tilex, tiley, samples, seed, sample_base, out_l, out_e, stream=None):
        """Weighted delta tracking against the scene's reference densities
        (Scene.set_reference_fog) into torch tensors (asynchronous)."""
        for t in (out_l, out_e):
            assert t.is_cuda and t.dtype.is_floating_point and t.is_contiguous() and t.numel() >= tilex * tiley * 4
        s = None if stream is None else C.c_void_p(stream.cuda_stream)
        L.check(L.lib().drk_kernel_render_weighted_device(self._h, int(x0), int(y0), int(tilex), int(tiley),
                                                          int(samples), int(seed), int(sample_base),
                                                          C.c_void_p(out_l.data_ptr()), C.c_void_p(out_e.data_ptr()),
                                                          s))

    def set_adjoint_estimator(self, estimator):
        """DRK_ADJ_SEGMENT (default) or DRK_ADJ_NULL_COLLISION."""
        L.check(L.lib().drk_kernel_set_adjoint_estimator(self._h, int(estimator)))

    def adjoint(self, x0, y0, tilex, tiley, samples, seed, sample_base, adj_lights, adj_env):
        """dL/d(FOG brick values) for the render with these arguments, given
        dL/d(phb sums) of both buffers as (tiley, tilex, 4) arrays (host)."""
        al = np.ascontiguousarray(adj_lights, dtype=np.float32).reshape(-1)
        ae = np.ascontiguousarray(adj_env, dtype=np.float32).reshape(-1)
        assert al.size == tilex * tiley * 4 and ae.size == al.size
        grad = np.zeros(self.scene.leafx[1] * self.scene.lvoxc3, np.float32)
        L.check(L.lib().drk_kernel_adjoint(self._h, int(x0), int(y0), int(tilex), int(tiley), int(samples),
                                           int(seed), int(sample_base), L.fptr(al), L.fptr(ae), L.fptr(grad)))
        return grad

    def adjoint_device(self, x0, y0, tilex, tiley, samples, seed, sample_base, adj_l, adj_e, grad, stream=None):
        """Device form: torch tensors; accumulates into grad (asynchronous)."""
        for t in (adj_l, adj_e, grad):
            assert t.is_cuda and t.dtype.is_floating_point and t.is_contiguous()
        s = None if stream is None else C.c_void_p(stream.cuda_stream)
        L.check(L.lib().drk_kernel_adjoint_device(self._h, int(x0), int(y0), int(tilex), int(tiley), int(samples),
                                                  int(seed), int(sample_base), C.c_void_p(adj_l.data_ptr()),
                                                  C.c_void_p(adj_e.data_ptr()), C.c_void_p(grad.data_ptr()), s))

    def set_event_budget(self, nbytes=None):
        """Limit of the adjoint's gradient-event buffer in bytes (None: the
        default, a quarter of device memory; 0: float atomics only)."""
        L.check(L.lib().drk_kernel_set_event_budget(self._h, (1 << 64) - 1 if nbytes is None else int(nbytes)))

    def footprint(self):
        """Device memory held by this kernel, in bytes."""
        n = C.c_size_t()
        L.check(L.lib().drk_kernel_footprint(self._h, C.byref(n)))
        return int(n.value)

    def record_bytes(self, x0, y0, tilex, tiley, samples):
        """Size of the per-sample path record of one render call."""
        n = C.c_size_t()
        L.check(L.lib().drk_kernel_record_bytes(self._h, int(x0), int(y0), int(tilex), int(tiley), int(samples),
                                                C.byref(n)))
        return int(n.value)

    def render_record_device(self, x0, y0, tilex, tiley, samples, seed, sample_base, out_l, out_e, rec,
                             stream=None):
        """render_device that also writes the path record into rec (a CUDA
        tensor of at least record_bytes() bytes) for adjoint_record_device."""
        for t in (out_l, out_e):
            assert t.is_cuda and str(t.dtype) == "torch.float32" and t.is_contiguous() and t.numel() >= tilex * tiley * 4
        assert rec.is_cuda and rec.is_contiguous()
        s = None if stream is None else C.c_void_p(stream.cuda_stream)
        L.check(L.lib().drk_kernel_render_record_device(
            self._h, int(x0), int(y0), int(tilex), int(tiley), int(samples), int(seed), int(sample_base),
            C.c_void_p(out_l.data_ptr()), C.c_void_p(out_e.data_ptr()), C.c_void_p(rec.data_ptr()),
            rec.numel() * rec.element_size(), s))

    def adjoint_record_device(self, x0, y0, tilex, tiley, samples, seed, sample_base, adj_l, adj_e, grad, rec,
                              stream=None):
        """adjoint_device from the path record of the same render call."""
        for t in (adj_l, adj_e, grad):
            assert t.is_cuda and t.dtype.is_floating_point and t.is_contiguous()
        assert rec.is_cuda and rec.is_contiguous()
        s = None if stream is None else C.c_void_p(stream.cuda_stream)
        L.check(L.lib().drk_kernel_adjoint_record_device(
            self._h, int(x0), int(y0), int(tilex), int(tiley), int(samples), int(seed), int(sample_base),
            C.c_void_p(adj_l.data_ptr()), C.c_void_p(adj_e.data_ptr()), C.c_void_p(grad.data_ptr()),
            C.c_void_p(rec.data_ptr()), rec.numel() * rec.element_size(), s))

    def count(self, x0, y0, tilex, tiley, samples, seed, sample_base=0):
        """Node-record reads and brick lookups of one render (counting pass)."""
        c = (C.c_ulonglong * 2)()
        L.check(L.lib().drk_kernel_count(self._h, int(x0), int(y0), int(tilex), int(tiley), int(samples), int(seed),
                                         int(sample_base), c))
        return int(c[0]), int(c[1])

    def Destroy(self):
        """RenderKernel::Destroy (src/kernel.cpp:609-615)."""
        if self._h is not None:
            L.lib().drk_kernel_destroy(self._h)
            self._h = None
            self._keep = None

    def __del__(self):
        try:
            self.Destroy()
        except Exception:
            pass

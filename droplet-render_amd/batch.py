"""Batched forward renders: many (kernel, tile) jobs in one persistent launch.

`Batch(jobs)` takes `(kernel, x0, y0, tilex, tiley, samples, seed,
sample_base, out_lights, out_env)` tuples (outputs: CUDA float32 tensors of
tilex*tiley*4) and wraps drk_batch_*: `launch()` enqueues the batched kernel,
`launch_graph()` replays it as a captured hipGraph, `autotune()` picks the
grid and the work-unit granularity by timing a sweep.

`render_record_views(jobs)` and `adjoint_record_views(jobs, grad)` run the
path-record forward and its adjoint of several calls in one persistent
launch each (drk_views_*): jobs are `(kernel, x0, y0, tilex, tiley, samples,
seed, sample_base, lights, env, record)` with CUDA float32 tensors (the sums,
or their gradients for the adjoint) and the record tensor.
"""
import ctypes as C

from . import _lib as L


def _stream_ptr(stream):
    return None if stream is None else C.c_void_p(stream.cuda_stream)


class Batch:
    def __init__(self, jobs):
        arr = (L.BatchJob * len(jobs))()
        self._keep = []
        for i, job in enumerate(jobs):
            k, x0, y0, tx, ty, spp, seed, base, out_l, out_e = job[:10]
            sq_l, sq_e = (job[10], job[11]) if len(job) > 10 else (None, None)
            for t in (out_l, out_e, sq_l, sq_e):
                if t is not None:
                    assert t.is_cuda and str(t.dtype) == "torch.float32" and t.is_contiguous()
                    assert t.numel() >= tx * ty * 4
            arr[i] = L.BatchJob(k._h, int(x0), int(y0), int(tx), int(ty), int(spp), int(seed), int(base),
                                out_l.data_ptr(), out_e.data_ptr(), sq_l.data_ptr() if sq_l is not None else None,
                                sq_e.data_ptr() if sq_e is not None else None)
            self._keep.append(k)
        h = C.c_void_p()
        L.check(L.lib().drk_batch_create(arr, len(jobs), C.byref(h)))
        self._h = h
        self.jobs = list(jobs)

    def set_tuning(self, grid_blocks=0, unit_scale=1.0):
        """Resident workgroups (0: full occupancy) and work-unit granularity:
        unit_scale x 16 sample chunks per pixel of every job."""
        L.check(L.lib().drk_batch_set_tuning(self._h, int(grid_blocks), float(unit_scale)))

    def launch(self, stream=None):
        L.check(L.lib().drk_batch_launch(self._h, _stream_ptr(stream)))

    def launch_graph(self, stream=None):
        L.check(L.lib().drk_batch_launch_graph(self._h, _stream_ptr(stream)))

    def autotune(self, stream=None, reps=2):
        """Times a grid x work-unit sweep, keeps the fastest; returns the table (text)."""
        buf = C.create_string_buffer(4096)
        L.check(L.lib().drk_batch_autotune(self._h, _stream_ptr(stream), int(reps), buf, 4096))
        return buf.value.decode()

    def destroy(self):
        if self._h is not None:
            L.lib().drk_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def _view_jobs(jobs):
    arr = (L.ViewJob * len(jobs))()
    for i, job in enumerate(jobs):
        k, x0, y0, tx, ty, spp, seed, base, t_l, t_e, rec = job
        for t in (t_l, t_e):
            assert t.is_cuda and str(t.dtype) == "torch.float32" and t.is_contiguous()
            assert t.numel() >= tx * ty * 4
        assert rec.is_cuda and rec.is_contiguous()
        arr[i] = L.ViewJob(k._h, int(x0), int(y0), int(tx), int(ty), int(spp), int(seed), int(base),
                           t_l.data_ptr(), t_e.data_ptr(), rec.data_ptr(), rec.numel() * rec.element_size())
    return arr


def render_record_views(jobs, stream=None):
    """Record forward of every job, one launch (same sums and records as
    separate render_record_device calls)."""
    L.check(L.lib().drk_views_render_record_device(_view_jobs(jobs), len(jobs), _stream_ptr(stream)))


def adjoint_record_views(jobs, grad, stream=None):
    """Adjoint of every job into `grad` (a CUDA float32 tensor), one launch
    and one event reduction (none inside the scene's gradient window)."""
    assert grad.is_cuda and str(grad.dtype) == "torch.float32" and grad.is_contiguous()
    L.check(L.lib().drk_views_adjoint_record_device(_view_jobs(jobs), len(jobs), C.c_void_p(grad.data_ptr()),
                                                    _stream_ptr(stream)))

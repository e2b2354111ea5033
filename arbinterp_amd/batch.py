"""Many fields in one launch, and the RK4 particle push (GPU only).

FieldBatch(interps) groups interpolators of one kind (same ndim, mode and
precision) so that their queries run in one persistent launch; push_rk4
advances particles under a = -grad U of a 3D scalar interpolant.  Neither has
a counterpart in the reference, which queries one field per call and dropped
its particle integrator in v1.8.
"""
from __future__ import annotations

import ctypes

import torch

from . import _native as N
from .ARBInterp import _check_out, _check_positions, _stream_handle


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class FieldBatch:
    def __init__(self, interps):
        self.interps = list(interps)
        if not self.interps:
            raise ValueError("empty batch")
        arr = (ctypes.c_void_p * len(self.interps))(*[t._h.value for t in self.interps])
        h = ctypes.c_void_p()
        N.check(N.lib().arb_batch_create(arr, len(self.interps), _stream_handle(torch),
                                         ctypes.byref(h)), "batch")
        self._h = h
        for t in self.interps:
            t._info.table_built = 1
        t0 = self.interps[0]
        self.ndim = t0.NDIM
        self.has_comps = bool(t0._info.has_comps)
        self.has_norm = bool(t0._info.has_norm)
        self.precision = t0.precision

    def set_tuning(self, chunk=0, threads=0, blocks_per_cu=0):
        N.check(N.lib().arb_batch_set_tuning(self._h, int(chunk), int(threads),
                                             int(blocks_per_cu)), "batch tuning")

    def query(self, q, counts, out_f64=False, check=False, out=None):
        """q: CUDA float64 tensor (sum(counts) x >= d), field k's rows after
        field k-1's.  Returns (comps, norms, grads) tensors for all rows."""
        if len(counts) != len(self.interps) or any(int(x) < 0 for x in counts):
            raise ValueError("counts must hold one non-negative count per field")
        n = int(sum(counts))
        d = self.ndim
        _check_positions(torch, q, d, "FieldBatch.query")
        if q.shape[0] < n:
            raise ValueError("q has %d rows, the counts need %d" % (q.shape[0], n))
        dt = torch.float64 if (self.precision == "fp64" or out_f64) else torch.float32
        if out is None:
            comps = torch.empty((n, 3), dtype=dt, device="cuda") if self.has_comps else None
            norms = torch.empty((n,), dtype=dt, device="cuda") if self.has_norm else None
            grads = torch.empty((n, d), dtype=dt, device="cuda") if self.has_norm else None
        else:
            comps, norms, grads = out
            if self.has_comps:
                _check_out(torch, comps, "comps", (n, 3), dt)
            if self.has_norm:
                _check_out(torch, norms, "norms", (n,), dt)
                _check_out(torch, grads, "grads", (n, d), dt)
        c = (ctypes.c_int64 * len(counts))(*[int(x) for x in counts])
        flags = (N.ARB_Q_OUT_F64 if out_f64 else 0) | (N.ARB_Q_CHECK if check else 0)
        N.check(N.lib().arb_batch_query(self._h, _ptr(q), c, q.stride(0), _ptr(comps),
                                        _ptr(norms), _ptr(grads), flags, _stream_handle(torch)),
                "batch query")
        return comps, norms, grads

    def close(self):
        if getattr(self, "_h", None):
            N.lib().arb_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def push_rk4(interp, x, v, dt, nsteps, steps_per_launch=10, graph=True):
    """Advance particles in place: x, v are CUDA float64 tensors (n x 3);
    classical RK4 with a = -grad U (unit mass).  Particles leaving the
    interpolation box become NaN and stay NaN."""
    if x.dtype != torch.float64 or v.dtype != torch.float64 or not x.is_contiguous() \
            or not v.is_contiguous() or x.shape != v.shape or x.shape[1] != 3:
        raise ValueError("x and v must be contiguous CUDA float64 tensors of shape (n, 3)")
    N.check(N.lib().arb_push_rk4(interp._h, _ptr(x), _ptr(v), x.shape[0], float(dt), int(nsteps),
                                 int(steps_per_launch), 1 if graph else 0, _stream_handle(torch)),
            "push")
    interp._info.table_built = 1

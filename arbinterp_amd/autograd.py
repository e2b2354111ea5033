"""Differentiable tricubic / quadcubic interpolation for PyTorch.

`InterpFunction.apply(positions, values, interp)` returns (U, grad U) at the
positions for the interpolator's scalar interpolant (scalar mode, or |B| for
'norm'/'both'), with gradients flowing to

  * positions: gv * grad U + Hess U . gg (the analytic Hessian of the
    interpolant, evaluated by the backward kernel; for mixed-precision 3D
    fields from the node tiles, whose FP64 model carries the cancelling part
    of the Hessian);
  * values (the stored node values, canonical x-fastest order, shape
    (ninterp, nnodes) or (nnodes,) for one column): the adjoint of the linear
    map values -> (U, grad U), computed deterministically on the GPU
    (per-cell sums in query-index order, fixed-order gather per node).

The interpolant is C1 only: its Hessian jumps across cell faces, so finite
difference checks of the position gradient must stay away from the faces.

`DifferentiableField` bundles an interpolator with trainable node values.
"""
from __future__ import annotations

import ctypes

import torch

from . import _native as N
from .ARBInterp import _stream_handle


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class InterpFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, positions, values, interp):
        d = interp.NDIM
        if positions.dim() != 2 or positions.shape[1] < d:
            raise ValueError("positions must be (n, >= %d)" % d)
        q = positions[:, :d].detach().to(device="cuda", dtype=torch.float64).contiguous()
        vals = values.detach().to(device="cuda", dtype=torch.float64).contiguous()
        if vals.numel() != interp._info.ninterp * interp._info.nnodes:
            raise ValueError("values must hold ninterp x nnodes = %d x %d numbers"
                             % (interp._info.ninterp, interp._info.nnodes))
        N.check(N.lib().arb_field_set_values(interp._h, _ptr(vals), _stream_handle(torch)),
                "set_values")
        interp._info.table_built = 0
        # forward and backward read the same coefficient source: node tiles
        # for mixed 3D fields (an FP64 model per block plus FP32 residuals;
        # no table is built), the table otherwise; one build per set of values
        _, U, G = interp.query_device(q, out_f64=True)
        ctx.save_for_backward(q)
        ctx.interp = interp
        ctx.pos_dtype = positions.dtype
        ctx.pos_cols = positions.shape[1]
        ctx.val_shape = values.shape
        ctx.val_dtype = values.dtype
        return U.to(values.dtype), G.to(values.dtype)

    @staticmethod
    def backward(ctx, gU, gG):
        (q,) = ctx.saved_tensors
        interp = ctx.interp
        n, d = q.shape
        gv = (gU if gU is not None else torch.zeros(n, device="cuda")).to(torch.float64).contiguous()
        gg = (gG if gG is not None else torch.zeros(n, d, device="cuda")).to(torch.float64).contiguous()
        need_q, need_v = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        dq = torch.empty((n, d), dtype=torch.float64, device="cuda") if need_q else None
        nn = interp._info.nnodes
        dval = torch.empty((nn,), dtype=torch.float64, device="cuda") if need_v else None
        if need_q or need_v:
            N.check(N.lib().arb_query_backward(interp._h, _ptr(q), n, d, _ptr(gv), _ptr(gg),
                                               _ptr(dq), _ptr(dval), _stream_handle(torch)),
                    "backward")
        gq = None
        if need_q:
            gq = dq.to(ctx.pos_dtype)
            if ctx.pos_cols > d:
                gq = torch.cat([gq, torch.zeros((n, ctx.pos_cols - d), dtype=ctx.pos_dtype,
                                                device=gq.device)], dim=1)
        gvals = None
        if need_v:
            k = interp._info.ninterp
            full = torch.zeros((k, nn), dtype=torch.float64, device="cuda")
            full[k - 1] = dval
            gvals = full.reshape(ctx.val_shape).to(ctx.val_dtype)
        return gq, gvals, None


def node_values(interp):
    """The interpolator's stored node values as a (ninterp, nnodes) FP64 tensor."""
    out = torch.empty((interp._info.ninterp, interp._info.nnodes), dtype=torch.float64,
                      device="cuda")
    N.check(N.lib().arb_field_values(interp._h, _ptr(out), _stream_handle(torch)), "values")
    return out


class DifferentiableField(torch.nn.Module):
    """An interpolator whose node values are a trainable parameter.

    >>> fld = DifferentiableField(tricubic(rows, 'quiet'), dtype=torch.float32)
    >>> U, G = fld(positions)      # differentiable in positions and fld.values
    """

    def __init__(self, interp, dtype=torch.float64):
        super().__init__()
        self.interp = interp
        v = node_values(interp)
        if v.shape[0] == 1:
            v = v[0]
        self.values = torch.nn.Parameter(v.to(dtype))

    def forward(self, positions):
        return InterpFunction.apply(positions, self.values, self.interp)

"""Drop-in `tricubic` / `quadcubic` classes backed by gfx950 HIP kernels.

Same constructor and query API as the reference ARBInterp v1.8
(src/ARBInterp/ARBInterp.py): `tricubic(field, *args, **kwargs)` with the
positional 'quiet' switch and `mode='vector'|'norm'|'both'`, then `.Query(q)`,
`.rQuery(Q)`, `.sQuery(q)`, `.allCoeffs()`, and the public attributes
(nPos, hx/hy/hz[/ht], x/y/z[/t]IntMin/Max, nc, queryInds, basePointInds, A,
inputfield, eps).  Return shapes, NaN rows written into the caller's array,
banners, `SystemExit` on badly shaped fields and the IndexError -> sQuery
fallback follow the reference.

Extensions (keyword arguments the reference does not have; its defaults are
kept): `precision='fp64'` (default, plain FP64) or `'mixed'` (FP32 table and
evaluation), `table_limit=None` (bytes: a larger coefficient table is not
held and queries form coefficients on the fly, as they do when the table
does not fit in device memory), `quiet=True` as a keyword, `boundary=None` (default, the
reference's box) or `'neumann'` / `'dirichlet'` (one ghost node layer per
face, so the whole sampled box is interpolated; the attributes then describe
the extended grid), and `query_device(q)` for positions already in GPU memory
(a torch tensor), which returns torch tensors.

Every numeric result comes from libarbinterp.so; there is no CPU path.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as N

__version__ = "1.8"

_BANNERS = {
    "scalar": "--- Scalar field, ignoring switches, interpolating for magnitude and gradient --- \n",
    "vector": "--- Vector field, interpolating for vector components --- \n",
    "norm": "--- Vector field, interpolating for magnitude and gradient --- \n",
    "both": "--- Vector field, interpolating vector components plus magnitude and gradient --- \n",
    "invalid": "--- Vector field, invalid option, defaulting to interpolating for vector components --- \n",
    "default": "--- Vector field, no option selected, defaulting to interpolating for vector components --- \n",
}

_MODE_CODE = {"scalar": N.ARB_MODE_SCALAR, "vector": N.ARB_MODE_VECTOR, "norm": N.ARB_MODE_NORM,
              "both": N.ARB_MODE_BOTH, "invalid": N.ARB_MODE_VECTOR_NAN,
              "default": N.ARB_MODE_VECTOR_NAN}

_BC_CODE = {None: N.ARB_BC_NONE, "none": N.ARB_BC_NONE, "neumann": N.ARB_BC_NEUMANN,
            "dirichlet": N.ARB_BC_DIRICHLET}

_AX = "xyzt"


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("arbinterp_amd needs a ROCm GPU (gfx950); none is visible")
    return torch


def _stream_handle(torch):
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _check_out(torch, t, name, shape, dtype, align=8):
    """A caller-supplied output buffer must be a contiguous CUDA tensor of the
    given shape and dtype (the kernels write whole rows through its pointer);
    `align` is the byte alignment of its data (16 for packed records)."""
    if t is None:
        raise ValueError("out: %s buffer is missing" % name)
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise ValueError("out: %s must be a CUDA tensor" % name)
    if t.dtype != dtype:
        raise ValueError("out: %s must be %s, got %s" % (name, dtype, t.dtype))
    if tuple(t.shape) != tuple(shape):
        raise ValueError("out: %s must have shape %s, got %s" % (name, tuple(shape),
                                                                 tuple(t.shape)))
    if not t.is_contiguous():
        raise ValueError("out: %s must be contiguous" % name)
    if t.data_ptr() % align:
        raise ValueError("out: %s must be %d-byte aligned" % (name, align))


def _check_positions(torch, q, d, what):
    if not isinstance(q, torch.Tensor) or q.dtype != torch.float64 or not q.is_cuda \
            or q.dim() != 2 or q.shape[1] < d or (q.shape[0] > 1 and q.stride(1) != 1) \
            or q.stride(0) < d:
        raise ValueError("%s needs a CUDA float64 tensor (n x >= %d) with contiguous rows"
                         % (what, d))


def _pinned_copy(torch, x):
    """An asynchronous copy of device tensor x into new pinned host memory
    (None for None); the caller synchronises the stream before reading it."""
    if x is None:
        return None
    h = torch.empty(x.shape, dtype=x.dtype, pin_memory=True)
    h.copy_(x, non_blocking=True)
    return h


class _Interp:
    """Shared implementation of the 3D and 4D classes."""

    NDIM = 3
    __version__ = "1.8"

    def __init__(self, field, *args, **kwargs):
        d = self.NDIM
        self.eps = 10 * np.finfo(float).eps            # ARBInterp.py:11
        precision = kwargs.pop("precision", "fp64")
        boundary = kwargs.pop("boundary", None)
        table_limit = kwargs.pop("table_limit", None)
        quiet = ("quiet" in args) or bool(kwargs.pop("quiet", False))
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        if boundary not in _BC_CODE:
            raise ValueError("boundary must be None, 'neumann' or 'dirichlet'")
        self.precision = precision
        torch = _torch()
        is_tensor = isinstance(field, torch.Tensor)
        shape = tuple(field.shape)
        if len(shape) != 2:
            raise ValueError("field must be a 2-D array of rows")
        ncols = shape[1]
        # variant dispatch (ARBInterp.py:24-104, :643-723)
        if ncols == d + 1:
            kind = "scalar"
        elif ncols == d + 3:
            if "mode" in kwargs:
                m = kwargs["mode"]
                kind = m if m in ("vector", "norm", "both") else "invalid"
            else:
                kind = "default"
        else:
            raise SystemExit("--- Input not shaped as expected - should be N x %d or N x %d ---"
                             % (d + 1, d + 3))
        if not quiet:
            print(_BANNERS[kind])
        self._kind = kind
        self._variant = {"scalar": 2, "vector": 1, "norm": 2, "both": 3}.get(kind, 1)
        self._field_src = field
        # a float32 field queried with float32 positions reduces in float32
        # (numpy promotion in the reference, ARBInterp.py:358-373)
        self._field_f32 = (field.dtype == torch.float32) if is_tensor else \
            (np.asarray(field).dtype == np.float32)
        if is_tensor:
            rows = field.detach().to(device="cuda", dtype=torch.float64).contiguous()
        else:
            rows = torch.from_numpy(np.ascontiguousarray(field, dtype=np.float64)).to("cuda")
        L = N.lib()
        h = ctypes.c_void_p()
        prec = N.ARB_PREC_FP64 if precision == "fp64" else N.ARB_PREC_MIXED
        # a float32 array is normed in float32 (numpy.linalg.norm keeps its dtype)
        mcode = _MODE_CODE[kind] | (N.ARB_MODE_F32_NORM if self._field_f32 else 0)
        rc = L.arb_field_create(ctypes.c_void_p(rows.data_ptr()), rows.shape[0], ncols,
                                rows.shape[1], d, mcode, prec,
                                _stream_handle(torch), ctypes.byref(h))
        N.check(rc, "field")
        self._h = h
        if _BC_CODE[boundary] != N.ARB_BC_NONE:
            N.check(L.arb_field_set_boundary(h, _BC_CODE[boundary], _stream_handle(torch)),
                    "boundary")
        self.boundary = boundary
        if table_limit is not None:
            N.check(L.arb_set_table_limit(h, int(table_limit)), "table_limit")
        info = N.FieldInfo()
        N.check(L.arb_field_info_get(h, ctypes.byref(info)))
        self._info = info
        self.nc = int(info.nc)
        n = [int(info.n[a]) for a in range(d)]
        self._n = n
        self.nPos = np.array([n[a] - 3 for a in range(d)])         # ARBInterp.py:545
        for a in range(d):
            setattr(self, "h" + _AX[a], np.float64(info.h[a]))
            setattr(self, _AX[a] + "IntMin", np.float64(info.int_min[a]))
            setattr(self, _AX[a] + "IntMax", np.float64(info.int_max[a]))
        if d == 4:                                                  # ARBInterp.py:1288-1291
            self.nPosx, self.nPosy, self.nPosz, self.nPost = n
        self._lo = np.array([info.int_min[a] for a in range(d)])
        self._hi = np.array([info.int_max[a] for a in range(d)])
        self._hh = np.array([info.h[a] for a in range(d)])
        self.queryInds = None
        self.default_flags = 0      # query-path choice set by arbinterp_amd.autotune

    # -- reference attributes -------------------------------------------------
    @property
    def A(self):
        """A = inv(B).D (ARBInterp.py:175, :878), built exactly on the host."""
        K = 4 ** self.NDIM
        out = np.zeros((K, K))
        N.check(N.lib().arb_operator_matrix(self.NDIM, 3, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    @property
    def basePointInds(self):
        """Flat node ids of each cell's base node (ARBInterp.py:558-565)."""
        d = self.NDIM
        stride = [int(np.prod(self._n[:a])) for a in range(d)]
        base = np.zeros(1, dtype=np.int64) + sum(stride)
        for a in range(d):
            base = (base[None, :] + stride[a] * np.arange(self._n[a] - 3)[:, None]).reshape(-1)
        return np.sort(base)

    @property
    def inputfield(self):
        """The field rows in the reference's sorted order (ARBInterp.py:530-532)."""
        f = self._field_src
        if not isinstance(f, np.ndarray):
            f = f.detach().cpu().numpy()
        keys = tuple(f[:, a] for a in range(self.NDIM))   # lexsort: last key is primary
        return f[np.lexsort(keys)]

    @property
    def alphamask(self):
        m = np.zeros((self.nc + 1, 1))
        m[-1] = 1
        if self._info.table_built:
            m[:] = 1
        return m

    def allCoeffs(self):
        """Build the whole coefficient table on the GPU (ARBInterp.py:523-525).
        A table that does not fit in device memory (or in `table_limit`) is
        not built: the field then forms each bucket's coefficients on the fly
        in every query, with the same results, as the reference fills only
        the cells its queries touch (ARBInterp.py:376-377)."""
        torch = _torch()
        rc = N.lib().arb_build_coeffs(self._h, _stream_handle(torch))
        if rc == N.ARB_ENOMEM and self.field_info().on_the_fly:
            return
        N.check(rc, "allCoeffs")
        torch.cuda.current_stream().synchronize()
        self._info.table_built = 1

    def set_table_limit(self, max_bytes):
        """Largest coefficient table this field may hold (bytes; 0 = no limit
        but device memory).  Above it, queries form coefficients on the fly
        (bitwise the table's results); the calls that read the table itself
        (coefficient_table, the backward, batches, the particle push) raise
        MemoryError."""
        N.check(N.lib().arb_set_table_limit(self._h, int(max_bytes)), "table_limit")
        self._info = self.field_info()

    def timings(self):
        """Stage durations (ms) of the last query made with ARB_Q_TIMING:
        [keys, grouping or radix sort (or bucket scan), scatter, evaluation, split into
        arrays]; the direct path reports its kernel as the evaluation."""
        ms = (ctypes.c_float * 5)()
        N.check(N.lib().arb_last_timings(self._h, ms))
        return [float(x) for x in ms]

    def field_info(self):
        """The library's current description of the field (arb_field_info):
        grid constants, table state, bucket block, node-tile bytes."""
        info = N.FieldInfo()
        N.check(N.lib().arb_field_info_get(self._h, ctypes.byref(info)))
        return info

    def set_tuning(self, block_cells=None, threads=0):
        """Cells per bucket block (one entry per axis, None = default) and
        threads per evaluation workgroup for the bucketed query path."""
        arr = None
        if block_cells is not None:
            arr = (ctypes.c_int32 * self.NDIM)(*[int(b) for b in block_cells])
        N.check(N.lib().arb_set_tuning(self._h, arr, int(threads)), "tuning")

    def coefficient_table(self, cell0=0, ncells=None):
        """Coefficients of cells [cell0, cell0+ncells) as a (ncells, ninterp, K)
        array, interpolants ordered (x, y, z, n) as present; row nc is the
        sentinel (the reference's alphax/.../alphan[:, cell], ARBInterp.py:42-74)."""
        torch = _torch()
        if ncells is None:
            ncells = self.nc + 1 - cell0
        K = 4 ** self.NDIM
        dt = torch.float64 if self.precision == "fp64" else torch.float32
        out = torch.empty((ncells, self._info.ninterp, K), dtype=dt, device="cuda")
        N.check(N.lib().arb_table_copy(self._h, cell0, ncells, ctypes.c_void_p(out.data_ptr()),
                                       _stream_handle(torch)), "table")
        self._info.table_built = 1
        return out.cpu().numpy()

    def _table_meta(self):
        d = self.NDIM
        info = self._info
        return np.array([d, info.mode, info.precision, info.ninterp]
                        + [info.n[a] for a in range(d)], dtype=np.int64), \
            np.array([info.h[a] for a in range(d)] + [info.int_min[a] for a in range(d)]
                     + [info.int_max[a] for a in range(d)], dtype=np.float64)

    def save_coefficients(self, path):
        """Write the whole coefficient table (built if needed) with the grid
        constants it belongs to, as an .npz file."""
        imeta, fmeta = self._table_meta()
        np.savez(path, table=self.coefficient_table(), imeta=imeta, fmeta=fmeta,
                 version=np.array(__version__))

    def load_coefficients(self, path):
        """Load a table written by save_coefficients for the same grid, mode and
        precision (checked exactly); the field's own table is not rebuilt."""
        torch = _torch()
        z = np.load(path, allow_pickle=False)
        imeta, fmeta = self._table_meta()
        if not (np.array_equal(z["imeta"], imeta) and np.array_equal(z["fmeta"], fmeta)):
            raise ValueError("coefficient file belongs to a different grid, mode or precision")
        tab = z["table"]
        dt = torch.float64 if self.precision == "fp64" else torch.float32
        src = torch.from_numpy(np.ascontiguousarray(tab)).to(device="cuda", dtype=dt)
        if src.shape[0] != self.nc + 1:
            raise ValueError("coefficient file has %d cells, field has %d" % (src.shape[0] - 1, self.nc))
        N.check(N.lib().arb_table_load(self._h, 0, self.nc + 1, ctypes.c_void_p(src.data_ptr()),
                                       _stream_handle(torch)), "load")
        torch.cuda.current_stream().synchronize()
        self._info.table_built = 1

    # -- queries ---------------------------------------------------------------
    def _alloc_out(self, torch, n, dtype):
        d = self.NDIM
        dev = "cuda"
        comps = norms = grads = None
        if self._info.has_comps:
            comps = torch.empty((n, 3), dtype=dtype, device=dev)
        if self._info.has_norm:
            norms = torch.empty((n,), dtype=dtype, device=dev)
            grads = torch.empty((n, d), dtype=dtype, device=dev)
        return comps, norms, grads

    def _run(self, torch, q_dev, n, ldq, comps, norms, grads, cells, flags):
        ptr = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None
        rc = N.lib().arb_query(self._h, ptr(q_dev), n, ldq, ptr(comps), ptr(norms), ptr(grads),
                               ptr(cells), flags, _stream_handle(torch))
        self._info.table_built = 1
        return rc

    def query_device(self, q, nan_rows=False, out_f64=False, cells=False, flags=0, out=None):
        """Evaluate positions held in GPU memory (torch tensor n x >= d, float64,
        row-contiguous).  Returns (comps, norms, grads[, cells]) as torch tensors
        (None where the mode has no such output), enqueued on the current
        stream.  nan_rows=True writes NaN over out-of-range rows of `q`
        (ARBInterp.py:350-355).
        `out` = a tuple returned by an earlier call of the same size reuses
        those buffers (no allocation)."""
        torch = _torch()
        _check_positions(torch, q, self.NDIM, "query_device")
        n = q.shape[0]
        dt = torch.float64 if (self.precision == "fp64" or out_f64) else torch.float32
        if out is not None:
            if len(out) < (4 if cells else 3):
                raise ValueError("out must hold (comps, norms, grads%s)"
                                 % (", cells" if cells else ""))
            comps, norms, grads = out[:3]
            cel = out[3] if cells else None
            if self._info.has_comps:
                _check_out(torch, comps, "comps", (n, 3), dt)
            if self._info.has_norm:
                _check_out(torch, norms, "norms", (n,), dt)
                _check_out(torch, grads, "grads", (n, self.NDIM), dt)
            if cells:
                _check_out(torch, cel, "cells", (n,), torch.int64)
        else:
            comps, norms, grads = self._alloc_out(torch, n, dt)
            cel = torch.empty((n,), dtype=torch.int64, device="cuda") if cells else None
        if not flags & (N.ARB_Q_DIRECT | N.ARB_Q_SORTED | N.ARB_Q_SORTED_GLOBAL):
            flags |= self.default_flags
        fl = flags | (N.ARB_Q_NAN_ROWS if nan_rows else 0)
        if out_f64:
            fl |= N.ARB_Q_OUT_F64
        rc = self._run(torch, q, n, q.stride(0), comps, norms, grads, cel, fl)
        N.check(rc, "query")
        return (comps, norms, grads, cel) if cells else (comps, norms, grads)

    def query_packed(self, q, out=None, flags=0, out_f64=False):
        """Like query_device, but each row's outputs are written as one aligned
        record of `packed_width` elements, [comps(3)] [value, grad(d)], zero
        padded.  Returns the (n, packed_width) tensor; `unpack` gives views."""
        torch = _torch()
        _check_positions(torch, q, self.NDIM, "query_packed")
        n = q.shape[0]
        W = int(self._info.packed_width)
        dt = torch.float64 if (self.precision == "fp64" or out_f64) else torch.float32
        if out is not None:
            _check_out(torch, out, "records", (n, W), dt, align=16)
        rec = out if out is not None else torch.empty((n, W), dtype=dt, device="cuda")
        if not flags & (N.ARB_Q_DIRECT | N.ARB_Q_SORTED | N.ARB_Q_SORTED_GLOBAL):
            flags |= self.default_flags
        fl = flags | N.ARB_Q_PACKED | (N.ARB_Q_OUT_F64 if out_f64 else 0)
        N.check(self._run(torch, q, n, q.stride(0), rec, None, None, None, fl), "query")
        return rec

    def unpack(self, rec):
        """(comps, values, grads) views of a packed record tensor (None where
        the mode has no such output)."""
        d = self.NDIM
        comps = rec[:, 0:3] if self._info.has_comps else None
        vals = grads = None
        if self._info.has_norm:
            b = 3 if self._info.has_comps else 0
            vals = rec[:, b]
            grads = rec[:, b + 1:b + 1 + d]
        return comps, vals, grads

    def rQuery(self, query):
        """Batch query (rQuery1/2/3, ARBInterp.py:344-521 / :1064-1258).

        Rows with a coordinate strictly outside [IntMin, IntMax] are set to NaN
        in the caller's array (all columns); their outputs are NaN.  Sets
        `queryInds`.  Raises IndexError when a row reduces past the sentinel
        cell (the reference's alphamask lookup does)."""
        d = self.NDIM
        torch = _torch()
        n = len(query)
        qh = np.ascontiguousarray(query[:, :d], dtype=np.float64)
        q_dev = torch.from_numpy(qh).to("cuda", non_blocking=False)
        # whole caller rows with a coordinate strictly outside the box (the
        # same comparisons as numpy's on the caller's values, made on the
        # device copy; float32 values widen exactly); the kernels evaluate
        # such rows from the sentinel cell whether or not they hold NaN
        lo = torch.tensor([float(self._lo[a]) for a in range(d)], dtype=torch.float64,
                          device="cuda")
        hi = torch.tensor([float(self._hi[a]) for a in range(d)], dtype=torch.float64,
                          device="cuda")
        oob = ((q_dev < lo) | (q_dev > hi)).any(dim=1)
        out_dt = torch.float64
        comps, norms, grads = self._alloc_out(torch, n, out_dt)
        cells = torch.empty((n,), dtype=torch.int64, device="cuda")
        flags = (N.ARB_Q_CHECK | (N.ARB_Q_OUT_F64 if self.precision == "mixed" else 0)
                 | self.default_flags)
        if self._field_f32 and query.dtype == np.float32:
            flags |= N.ARB_Q_REDUCE_F32
        rc = self._run(torch, q_dev, n, d, comps, norms, grads, cells, flags)
        # results come back through pinned host buffers (one synchronisation)
        host = [_pinned_copy(torch, x) for x in (comps, norms, grads, cells)]
        nbad = int(oob.sum())
        bad = torch.nonzero(oob).squeeze(1).cpu().numpy() if nbad else None
        torch.cuda.current_stream().synchronize()
        if bad is not None:
            query[bad] = np.nan
        self.queryInds = host[3].numpy()
        if rc == N.ARB_EINDEX:
            raise IndexError("index %d is out of bounds for axis 0 with size %d"
                             % (int(self.queryInds.max()), self.nc + 1))
        N.check(rc, "query")
        out_c = host[0].numpy() if comps is not None else None
        out_n = host[1].numpy().reshape(n, 1) if norms is not None else None
        out_g = host[2].numpy() if grads is not None else None
        if self._variant == 1:
            return out_c
        if self._variant == 2:
            return out_n, out_g
        return out_c, out_n, out_g

    def sQuery(self, query):
        """Single-point query (sQuery1/2/3, ARBInterp.py:213-342 / :916-1062).
        Out of the box -> bare NaN; a NaN coordinate or a cell index past the
        sentinel raises IndexError; a 2-D argument makes the bounds test raise
        ValueError, as in the reference."""
        d = self.NDIM
        for a in range(d):
            if query[a] < self._lo[a] or query[a] > self._hi[a]:
                return np.nan
        pt = np.array([float(query[a]) for a in range(d)], dtype=np.float64)
        if np.isnan(pt).any():
            raise IndexError("index -9223372036854775808 is out of bounds for axis 0 with size %d"
                             % (self.nc + 1))
        torch = _torch()
        if self._field_f32 and np.asarray(query).dtype == np.float32:
            return self._point_f32(torch, pt)
        if getattr(self, "_pt_out", None) is None:
            # comps(3), norm, grad(4) FP64 and the cell, reused by every call
            self._pt_out = np.zeros(8, dtype=np.float64)
            self._pt_cell = ctypes.c_int64(0)
            base = self._pt_out.ctypes.data
            self._pt_ptrs = (ctypes.c_void_p(base), ctypes.c_void_p(base + 24),
                             ctypes.c_void_p(base + 32),
                             ctypes.c_void_p(ctypes.addressof(self._pt_cell)))
        cp, npp, gp, cellp = self._pt_ptrs
        rc = N.lib().arb_query_point(self._h, ctypes.c_void_p(pt.ctypes.data), cp, npp, gp,
                                     cellp, _stream_handle(torch))
        self._info.table_built = 1
        self.queryInd = int(self._pt_cell.value)
        if rc == N.ARB_EINDEX:
            raise IndexError("index %d is out of bounds for axis 0 with size %d"
                             % (self.queryInd, self.nc + 1))
        N.check(rc, "query")
        out = self._pt_out
        if self._variant == 1:
            return out[0:3].copy()
        nv = np.float64(out[3])
        gv = out[4:4 + d].copy()
        if self._variant == 2:
            return nv, gv
        return out[0:3].copy(), nv, gv

    def _point_f32(self, torch, pt):
        """A float32 point of a float32 field: the float32 reduction of the
        batch path on one row (the reference's sQuery reduces in float32 too;
        it also rounds the basis products to float32, which this path does
        not, a difference of order 1e-7 relative)."""
        d = self.NDIM
        q_dev = torch.from_numpy(pt[None, :]).to("cuda")
        comps, norms, grads = self._alloc_out(torch, 1, torch.float64)
        cells = torch.empty((1,), dtype=torch.int64, device="cuda")
        flags = N.ARB_Q_CHECK | N.ARB_Q_REDUCE_F32 | (N.ARB_Q_OUT_F64 if self.precision == "mixed"
                                                      else 0)
        rc = self._run(torch, q_dev, 1, d, comps, norms, grads, cells, flags)
        self.queryInd = int(cells.cpu().numpy()[0])
        if rc == N.ARB_EINDEX:
            raise IndexError("index %d is out of bounds for axis 0 with size %d"
                             % (self.queryInd, self.nc + 1))
        N.check(rc, "query")
        c = comps.cpu().numpy()[0] if comps is not None else None
        if self._variant == 1:
            return c
        nv = np.float64(norms.cpu().numpy()[0])
        gv = grads.cpu().numpy()[0]
        if self._variant == 2:
            return nv, gv
        return c, nv, gv

    def Query(self, query):
        """Query1/2/3 (ARBInterp.py:177-211 / :880-914): 2-D input with more
        than one column -> batch, else single point; an IndexError in the batch
        path is retried as a single-point query."""
        try:
            if query.shape[1] > 1:
                return self._unpack(self.rQuery(query))
            return self._unpack(self.sQuery(query))
        except IndexError:
            return self._unpack(self.sQuery(query))

    def _unpack(self, r):
        """Query2/Query3 unpack the result into 2 / 3 names, so a bare NaN
        from an out-of-box single point raises TypeError (ARBInterp.py:195,210)."""
        v = self._variant if hasattr(self, "_variant") else self.variant
        if v == 1:
            return r
        if v == 2:
            norm, grad = r
            return norm, grad
        comps, norm, grad = r
        return comps, norm, grad

    def close(self):
        if getattr(self, "_h", None):
            N.lib().arb_field_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class tricubic(_Interp):
    """Tricubic interpolation of a 3D field (ARBInterp.py:8-621)."""
    NDIM = 3


class quadcubic(_Interp):
    """Quadcubic interpolation of a 4D field (ARBInterp.py:627-1381)."""
    NDIM = 4

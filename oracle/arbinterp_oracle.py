"""CPU oracle: a numpy FP64 restatement of ARBInterp's Lekien-Marsden tricubic
(3D) and quadcubic (4D) interpolation, `src/ARBInterp/ARBInterp.py` v1.8.

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP product path:
only `tests/`, `__graft_entry__.smoke()` and the `cpu_baseline` leg of
`bench.py` import it.  The product package `arbinterp_amd` never imports it and
has no CPU fallback.

The restatement keeps the reference's algorithm and its vectorised numpy form
(N x 64 / N x 256 tile/repeat bases, column gathers from a (K, nc+1) table,
`sum(axis=1)`), so that it can also serve as the timed CPU baseline.  Every
function cites the reference line(s) it follows ("ARBInterp.py:N" is
`src/ARBInterp/ARBInterp.py` line N of the reference).

Pinning (each function is checked by tests that do not retype its formula):
  * make_B        -- sha256 of the reference's examples/B_Matrix_{3D,4D}.csv
                     values, plus the CSV itself when the reference is mounted.
  * make_D/make_A -- exact integer inverse / dyadic A; polynomial reproduction
                     (triquadratic exact); node interpolation = central
                     differences; an independent tensor-product cubic Hermite
                     construction of the interpolant; 4D known-answer test for
                     the row-shift quirk (f = xyzt -> f - hxhyhzht*H(u)H(v)H(w)H(t),
                     H(s) = s(1-s)^2); golden vectors of the reference.
  * field_params  -- shuffle invariance; closed-form base indices; golden.
  * cell reduction/query -- edge battery (IntMax aliasing, NaN rows),
                     golden vectors written by tests/golden/make_golden.py.
  * pad_nodes/extend_boundary (no reference counterpart) -- exact ghosts for
                     fields even / odd about every face keep the edge cells
                     as accurate as the interior (the other rule does not);
                     Neumann zero normal slope; Dirichlet reproduces linear
                     fields; interior unchanged; brute-force ghost indices.
"""
from __future__ import annotations

import itertools
import numpy as np

__all__ = [
    "make_B", "make_D", "make_A", "deriv_blocks", "corner_centres",
    "FieldParams", "field_params", "neighbour_offsets", "tricubic", "quadcubic",
    "coefficients_for_cells", "base_point_inds", "pad_nodes", "extend_boundary",
]


# --------------------------------------------------------------------------
# Operator matrices (ARBInterp.py:107-175 for 3D, :726-878 for 4D)
# --------------------------------------------------------------------------

def deriv_blocks(ndim: int):
    """Order of the b-vector blocks: f, then first derivatives, then mixed ones.

    3D (ARBInterp.py:118-125): f, fx, fy, fz, fxy, fxz, fyz, fxyz.
    4D (ARBInterp.py:738-757): f, fx, fy, fz, ft, fxy, fxz, fxt, fyz, fyt, fzt,
    fxyz, fxyt, fxzt, fyzt, fxyzt.  Each block is a tuple of axis ids.
    """
    blocks = []
    for order in range(ndim + 1):
        blocks.extend(itertools.combinations(range(ndim), order))
    return blocks


def _corners(ndim: int):
    # corner k = i + 2j + 4k' (+ 8l), x fastest (ARBInterp.py:110, :729)
    return [tuple((k >> a) & 1 for a in range(ndim)) for k in range(1 << ndim)]


def _exponents(ndim: int):
    # coefficient index i = ex + 4ey + 16ez (+ 64et) (ARBInterp.py:111, :730)
    return [tuple((i >> (2 * a)) & 3 for a in range(ndim)) for i in range(4 ** ndim)]


def make_B(ndim: int) -> np.ndarray:
    """Forward matrix: monomials and their mixed derivatives at the unit-cell
    corners (ARBInterp.py:110-125, :729-757).  Row = block*ncorner + corner,
    column = monomial index."""
    K = 4 ** ndim
    nco = 1 << ndim
    B = np.zeros((K, K), dtype=np.float64)
    corners = _corners(ndim)
    for i, ex in enumerate(_exponents(ndim)):
        for b, blk in enumerate(deriv_blocks(ndim)):
            for k, c in enumerate(corners):
                v = 1.0
                for a in range(ndim):
                    e, x = ex[a], float(c[a])
                    if a in blk:       # d/dx x^e = e x^(e-1)
                        v *= 0.0 if e == 0 else e * x ** (e - 1)
                    else:
                        v *= x ** e
                B[b * nco + k, i] = v
    return B


def corner_centres(ndim: int) -> np.ndarray:
    """Neighbourhood index of each cell corner inside the 4^d stencil block,
    offsets (1..2)^d: C = (21,22,25,26,37,38,41,42) in 3D (ARBInterp.py:129),
    (85,...,170) in 4D (ARBInterp.py:762)."""
    return np.array([sum((1 + c[a]) * 4 ** a for a in range(ndim))
                     for c in _corners(ndim)], dtype=np.int64)


def make_D(ndim: int) -> np.ndarray:
    """Finite-difference matrix in cuboid units (ARBInterp.py:129-173,
    :762-876): each block applies the tensor product of the central difference
    (-1/2, 0, +1/2) along the block's axes, centred at each corner.

    4D quirk reproduced (ARBInterp.py:860): the d4/dxdydzdt block is written by
    `enumerate(range(241,256))`, so row 240 stays zero and row 240+k (k>=1)
    carries the stencil centred at corner k-1.
    """
    K = 4 ** ndim
    nco = 1 << ndim
    C = corner_centres(ndim)
    D = np.zeros((K, K), dtype=np.float64)
    for b, blk in enumerate(deriv_blocks(ndim)):
        taps = [(0, 1.0)]
        for a in blk:
            taps = [(o + s * 4 ** a, w * 0.5 * s) for (o, w) in taps for s in (-1, 1)]
        rows = range(b * nco, (b + 1) * nco)
        centres = list(C)
        if ndim == 4 and len(blk) == 4:
            rows = range(b * nco + 1, (b + 1) * nco)      # 241..255
            centres = list(C[: nco - 1])                    # C[0..14]
        for r, c in zip(rows, centres):
            for o, w in taps:
                D[r, c + o] = w
    return D


def make_A(ndim: int) -> np.ndarray:
    """A = inv(B) @ D (ARBInterp.py:175, :878)."""
    return np.matmul(np.linalg.inv(make_B(ndim)), make_D(ndim))


_A_CACHE: dict = {}


def _A(ndim: int) -> np.ndarray:
    if ndim not in _A_CACHE:
        _A_CACHE[ndim] = make_A(ndim)
    return _A_CACHE[ndim]


# --------------------------------------------------------------------------
# Grid analysis (ARBInterp.py:528-568, :1264-1320)
# --------------------------------------------------------------------------

class FieldParams:
    """Sorted field and grid constants exactly as the reference derives them."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


def field_params(field: np.ndarray, ndim: int) -> FieldParams:
    """getFieldParams (ARBInterp.py:528-568 / :1264-1320).

    Sort: argsort of column 0, then stable argsorts of columns 1..ndim-1, so
    the order is last-axis major, x fastest.  Axis lengths are counted along
    the axis through the first node; h = |node0 - node1| per axis; IntMin /
    IntMax are the 2nd / 2nd-last node value per axis.
    """
    f = field[field[:, 0].argsort()]
    for a in range(1, ndim):
        f = f[f[:, a].argsort(kind="mergesort")]
    n = []
    for a in range(ndim):
        sel = np.ones(len(f), dtype=bool)
        for b in range(ndim):
            if b != a:
                sel &= f[:, b] == f[0, b]
        n.append(int(sel.sum()))
    axes = []
    for a in range(ndim):
        sel = np.ones(len(f), dtype=bool)
        for b in range(ndim):
            if b != a:
                sel &= f[:, b] == f[0, b]
        axes.append(f[sel, a])
    h = [np.abs(axes[a][0] - axes[a][1]) for a in range(ndim)]
    stride = [int(np.prod(n[:a])) for a in range(ndim)]
    int_min = [f[stride[a], a] for a in range(ndim)]
    int_max = [f[-2 * stride[a], a] for a in range(ndim)]
    base = base_point_inds(n, ndim)
    npos = [n[a] - 3 for a in range(ndim)]
    return FieldParams(field=f, n=n, npos=npos, h=h, int_min=int_min,
                       int_max=int_max, stride=stride, base=base, nc=len(base),
                       ndim=ndim)


def base_point_inds(n, ndim: int) -> np.ndarray:
    """basePointInds (ARBInterp.py:559-565): node ids (i_a in [1, n_a-3]) in
    canonical order."""
    stride = [int(np.prod(n[:a])) for a in range(ndim)]
    base = np.zeros(1, dtype=np.int64) + sum(stride)
    for a in range(ndim):
        base = (base[None, :] + stride[a] * np.arange(n[a] - 3)[:, None]).reshape(-1)
    return np.sort(base)


def neighbour_offsets(n, ndim: int) -> np.ndarray:
    """neighbourInd (ARBInterp.py:605-621 / :1357-1381) as offsets from a cell's
    base node: the 4^d block starting at (-1,...,-1), x fastest."""
    stride = [int(np.prod(n[:a])) for a in range(ndim)]
    off = np.zeros(1, dtype=np.int64) - sum(stride)
    for a in range(ndim):
        off = (off[None, :] + stride[a] * np.arange(4)[:, None]).reshape(-1)
    return off


def coefficients_for_cells(p: FieldParams, values: np.ndarray, cells: np.ndarray) -> np.ndarray:
    """calcCoefficients* (ARBInterp.py:570-603 / :1322-1355) for many cells:
    alpha[:, c] = A . F[neighbourInd(basePointInds[c])]   -> (K, len(cells))."""
    A = _A(p.ndim)
    off = neighbour_offsets(p.n, p.ndim)
    nodes = p.base[cells][None, :] + off[:, None]
    return A @ values[nodes]


# --------------------------------------------------------------------------
# Interpolators (drop-in mirror of the reference classes, for tests)
# --------------------------------------------------------------------------

_MODE_BANNERS = {
    "scalar": "--- Scalar field, ignoring switches, interpolating for magnitude and gradient --- \n",
    "vector": "--- Vector field, interpolating for vector components --- \n",
    "norm": "--- Vector field, interpolating for magnitude and gradient --- \n",
    "both": "--- Vector field, interpolating vector components plus magnitude and gradient --- \n",
    "invalid": "--- Vector field, invalid option, defaulting to interpolating for vector components --- \n",
    "default": "--- Vector field, no option selected, defaulting to interpolating for vector components --- \n",
}


class _Interp:
    NDIM = 3
    CHUNK = 1 << 16

    def __init__(self, field, *args, **kwargs):
        d = self.NDIM
        p = field_params(field, d)
        self.p = p
        self.inputfield = p.field
        self.nc = p.nc
        self.basePointInds = p.base
        self.A = _A(d)
        K = 4 ** d
        ncols = field.shape[1]
        quiet = "quiet" in args
        # variant dispatch (ARBInterp.py:24-104 / :643-723)
        if ncols == d + 1:
            kind, variant = "scalar", 2
        elif ncols == d + 3:
            if "mode" in kwargs:
                m = kwargs["mode"]
                kind = m if m in ("vector", "norm", "both") else "invalid"
            else:
                kind = "default"
            variant = {"vector": 1, "norm": 2, "both": 3}.get(kind, 1)
        else:
            raise SystemExit("--- Input not shaped as expected - should be N x %d or N x %d ---"
                             % (d + 1, d + 3))
        if not quiet:
            print(_MODE_BANNERS[kind])
        self.kind, self.variant = kind, variant
        f = p.field
        self.interps = {}
        # sentinel column nc (_init_tables): NaN except scalar mode (zero) and
        # explicit 'vector' mode (zero at nc, NaN at nc+1)
        # (ARBInterp.py:31,42-45,57,70,85,99)
        if variant in (1, 3):
            for j, name in enumerate("xyz"):
                self.interps[name] = f[:, d + j]
        if variant in (2, 3):
            if kind == "scalar":
                self.interps["n"] = f[:, d]
            else:
                self.interps["n"] = np.linalg.norm(f[:, d:], axis=1)
        self._init_tables()

    def _init_tables(self):
        """Coefficient tables (K, nc+1[+1]) with the sentinel columns, and the
        fill mask."""
        K = 4 ** self.NDIM
        kind = self.kind
        ncol_tab = self.nc + 2 if kind == "vector" else self.nc + 1
        sentinel = 0.0 if kind in ("scalar", "vector") else np.nan
        self.tables = {}
        for k in self.interps:
            t = np.zeros((K, ncol_tab))
            t[:, self.nc] = sentinel
            if kind == "vector":
                t[:, self.nc + 1] = np.nan
            self.tables[k] = t
        self.alphamask = np.zeros(self.nc + 1, dtype=bool)
        self.alphamask[-1] = True

    # -- coefficients -------------------------------------------------------
    def _fill(self, cells):
        cells = np.unique(cells)
        cells = cells[~self.alphamask[cells]]
        if len(cells) == 0:
            return
        for k, vals in self.interps.items():
            self.tables[k][:, cells] = coefficients_for_cells(self.p, vals, cells)
        self.alphamask[cells] = True

    def condition_tables(self, cells):
        """|A| . |F[nbr]| per interpolant for the given cells: the bound on
        sum_j |A_ij f_j| that FP64 rounding of alpha_i scales with."""
        A = np.abs(_A(self.NDIM))
        off = neighbour_offsets(self.p.n, self.NDIM)
        nodes = self.p.base[cells][None, :] + off[:, None]
        return {k: A @ np.abs(v[nodes]) for k, v in self.interps.items()}

    def condition_scale(self, query):
        """Per-output condition scale S = sum_i |beta_i| sum_j |A_ij f_j| with
        beta the value basis (or d basis / h for gradients), for the rows of
        `query` (a copy is reduced; NaN where the row is out of range).
        Any FP64 evaluation of the interpolant has an error of order eps*S."""
        d = self.NDIM
        inds, local = self.cell_reduce(np.array(query, dtype=np.float64, copy=True))
        ok = inds < self.nc
        N = len(inds)
        res = {k: np.full(N, np.nan) for k in self.interps}
        gres = np.full((N, d), np.nan)
        if ok.any():
            cells = inds[ok]
            m, dm = self._basis(local[ok])
            prod = m[0]
            for a in range(1, d):
                prod = prod * m[a]
            ct = self.condition_tables(cells)
            for k in ct:
                res[k][ok] = (np.abs(prod) * ct[k].T).sum(axis=1)
            if "n" in ct:
                for a in range(d):
                    b = dm[a] if a == 0 else m[0]
                    for c in range(1, d):
                        b = b * (dm[c] if c == a else m[c])
                    gres[ok, a] = (np.abs(b) * ct["n"].T).sum(axis=1) / self.p.h[a]
        comps = np.stack([res[c] for c in "xyz"], axis=1) if "x" in res else None
        norms = res["n"][:, None] if "n" in res else None
        grads = gres if "n" in res else None
        if self.variant == 1:
            return comps
        if self.variant == 2:
            return norms, grads
        return comps, norms, grads

    # -- derivatives for the backward pass (no reference counterpart) --------
    def _second_basis(self, local, a, b):
        """d^2 m_i / du_a du_b for every coefficient i, tile/repeat form."""
        N = len(local)
        d = self.NDIM
        K = 4 ** d
        one, zero = np.ones(N), np.zeros(N)

        def expand(v, ax):
            return np.tile(np.repeat(v, 4 ** ax, axis=1), K // 4 ** (ax + 1))

        out = None
        for ax in range(d):
            u = local[:, ax]
            if ax == a and ax == b:
                v = np.transpose(np.array([zero, zero, 2 * one, 6 * u]))
            elif ax == a or ax == b:
                v = np.transpose(np.array([zero, one, 2 * u, 3 * u ** 2]))
            else:
                v = np.transpose(np.array([one, u, u ** 2, u ** 3]))
            e = expand(v, ax)
            out = e if out is None else out * e
        return out

    def hessian(self, query):
        """Hessian of the scalar interpolant (physical units), (N, d, d)."""
        d = self.NDIM
        inds, local = self.cell_reduce(np.array(query, dtype=np.float64, copy=True))
        self._fill(inds[inds < self.nc])
        t = np.transpose(self.tables["n"][:, inds])
        H = np.zeros((len(inds), d, d))
        for a in range(d):
            for b in range(a, d):
                v = (t * self._second_basis(local, a, b)).sum(axis=1) / (self.p.h[a] * self.p.h[b])
                H[:, a, b] = v
                H[:, b, a] = v
        return H

    def backward(self, query, gv, gg):
        """Adjoint of (value, gradient) of the scalar interpolant.

        dq[i]   = gv_i grad U(x_i) + Hess U(x_i) . gg_i
        dval[n] = sum_i sum_j c_ij A_{j, l(n)}, c_i = gv_i m(x_i) + sum_k gg_ik dm_k(x_i)/h_k,
        scattered onto the nodes of each cell's neighbourhood (np.add.at).
        Rows outside the box contribute nothing.  dval is in canonical node
        order (the sorted field's row order)."""
        d = self.NDIM
        q = np.array(query, dtype=np.float64, copy=True)
        inds, local = self.cell_reduce(q.copy())
        ok = inds < self.nc
        N = len(inds)
        gv = np.asarray(gv, dtype=np.float64).reshape(N)
        gg = np.asarray(gg, dtype=np.float64).reshape(N, d)
        self._fill(inds[ok])
        t = np.transpose(self.tables["n"][:, inds[ok]])
        m, dm = self._basis(local[ok])
        prod = m[0]
        for a in range(1, d):
            prod = prod * m[a]
        grads = []
        dbas = []
        for a in range(d):
            b = dm[a] if a == 0 else m[0]
            for c in range(1, d):
                b = b * (dm[c] if c == a else m[c])
            dbas.append(b / self.p.h[a])
            grads.append((t * dbas[-1]).sum(axis=1))
        G = np.stack(grads, axis=1)
        H = self.hessian(q[ok])
        dq = np.full((N, d), np.nan)
        dq[ok] = gv[ok, None] * G + np.einsum("nab,nb->na", H, gg[ok])
        coef = gv[ok, None] * prod
        for a in range(d):
            coef = coef + gg[ok, a:a + 1] * dbas[a]
        A = _A(d)
        bbar = coef @ A                                    # (n_ok, K): sum_j c_j A_jl
        off = neighbour_offsets(self.p.n, d)
        nodes = self.p.base[inds[ok]][:, None] + off[None, :]
        dval = np.zeros(len(self.p.field))
        np.add.at(dval, nodes.reshape(-1), bbar.reshape(-1))
        return dq, dval

    def allCoeffs(self):
        """ARBInterp.py:523-525 / :1260-1262."""
        step = 1 << 15
        for s in range(0, self.nc, step):
            self._fill(np.arange(s, min(self.nc, s + step)))

    # -- cell reduction (ARBInterp.py:350-373 / :1069-1092) ------------------
    def cell_reduce(self, query):
        """Bounds mask (written into the caller's rows, ARBInterp.py:350-355),
        cuboid coordinates, float flat index with NaN -> nc, local
        coordinates."""
        p = self.p
        d = self.NDIM
        for a in range(d):
            query[np.where(query[:, a] < p.int_min[a])[0]] = np.nan
            query[np.where(query[:, a] > p.int_max[a])[0]] = np.nan
        iu = [(query[:, a] - p.int_min[a]) / p.h[a] for a in range(d)]
        ix = [np.floor(v) for v in iu]
        if d == 3:
            flat = ix[0] + ix[1] * (p.npos[0]) + ix[2] * (p.npos[0]) * (p.npos[1])
        else:
            flat = (ix[0] + ix[1] * (p.npos[0]) + ix[2] * (p.npos[0]) * (p.npos[1])
                    + ix[3] * (p.npos[0]) * (p.npos[1]) * (p.npos[2]))
        flat[np.where(np.isnan(flat))] = self.nc
        inds = flat.astype(int)
        local = np.stack([iu[a] - ix[a] for a in range(d)], axis=1)
        return inds, local

    # -- batch evaluation (ARBInterp.py:380-397,435-454,493-521) --------------
    def _basis(self, local):
        """Tile/repeat bases exactly as the reference builds them."""
        N = len(local)
        d = self.NDIM
        one, zero = np.ones(N), np.zeros(N)
        vec = [np.transpose(np.array([one, local[:, a], local[:, a] ** 2, local[:, a] ** 3]))
               for a in range(d)]
        dvec = [np.transpose(np.array([zero, one, 2 * local[:, a], 3 * local[:, a] ** 2]))
                for a in range(d)]
        K = 4 ** d

        def expand(v, a):
            # axis a varies with stride 4^a inside the K-wide basis
            return np.tile(np.repeat(v, 4 ** a, axis=1), K // 4 ** (a + 1))

        return [expand(vec[a], a) for a in range(d)], [expand(dvec[a], a) for a in range(d)]

    def _eval(self, inds, local):
        d = self.NDIM
        p = self.p
        N = len(inds)
        m, dm = self._basis(local)
        prod = m[0]
        for a in range(1, d):
            prod = prod * m[a]
        out = {}
        for k in self.tables:
            t = np.transpose(self.tables[k][:, inds])
            out[k] = (t * prod).sum(axis=1)
        grads = None
        if "n" in self.tables:
            t = np.transpose(self.tables["n"][:, inds])
            g = []
            for a in range(d):
                b = dm[a] if a == 0 else m[0]
                for c in range(1, d):
                    b = b * (dm[c] if c == a else m[c])
                g.append(((t * b) / p.h[a]).sum(axis=1))
            grads = np.transpose(np.array(g))
        comps = None
        if "x" in out:
            comps = np.hstack([np.reshape(out[c], (N, 1)) for c in "xyz"])
        norms = np.reshape(out["n"], (N, 1)) if "n" in out else None
        return comps, norms, grads

    def rQuery(self, query):
        """rQuery1/2/3 (ARBInterp.py:344-521 / :1064-1258)."""
        inds, local = self.cell_reduce(query)
        self.queryInds = inds
        if np.any(inds > self.nc):
            # alphamask[queryInds] raises IndexError (ARBInterp.py:376)
            raise IndexError("index out of bounds for alphamask")
        self._fill(inds[~self.alphamask[inds]])
        N = len(inds)
        parts = []
        for s in range(0, N, self.CHUNK):
            parts.append(self._eval(inds[s:s + self.CHUNK], local[s:s + self.CHUNK]))
        def cat(i):
            if parts and parts[0][i] is None:
                return None
            if not parts:
                return np.zeros((0, 3 if i == 0 else (1 if i == 1 else self.NDIM)))
            return np.concatenate([q[i] for q in parts], axis=0)
        comps, norms, grads = cat(0), cat(1), cat(2)
        if self.variant == 1:
            return comps
        if self.variant == 2:
            return norms, grads
        return comps, norms, grads

    def sQuery(self, query):
        """sQuery1/2/3 (ARBInterp.py:213-342 / :916-1062).  An out-of-box point
        returns a bare NaN; a 2-D argument makes numpy raise ValueError in the
        bounds test, as in the reference; a NaN coordinate or a flat index past
        the sentinel raises IndexError from the alphamask lookup."""
        p = self.p
        d = self.NDIM
        for a in range(d):
            if query[a] < p.int_min[a] or query[a] > p.int_max[a]:
                return np.nan
        iu = [(query[a] - p.int_min[a]) / p.h[a] for a in range(d)]
        ix = [np.floor(v) for v in iu]
        flat = ix[0] + ix[1] * (p.npos[0]) + ix[2] * (p.npos[0]) * (p.npos[1])
        if d == 4:
            flat = flat + ix[3] * (p.npos[0]) * (p.npos[1]) * (p.npos[2])
        if np.isnan(flat) or flat > self.nc:
            # NaN -> INT64_MIN or index > nc: alphamask[...] raises IndexError
            raise IndexError("index out of bounds for alphamask")
        ind = int(flat)
        self.queryInd = ind
        self._fill(np.array([ind]))
        local = np.array([[iu[a] - ix[a] for a in range(d)]], dtype=np.float64)
        comps, norms, grads = self._eval(np.array([ind]), local)
        if self.variant == 1:
            return comps[0]
        if self.variant == 2:
            return norms[0, 0], grads[0]
        return comps[0], norms[0, 0], grads[0]

    def Query(self, query):
        """Query1/2/3 dispatch (ARBInterp.py:177-211 / :880-914): an IndexError
        anywhere in the batch path is retried as a single-point query."""
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


class tricubic(_Interp):
    """Oracle mirror of `tricubic` (ARBInterp.py:8-621)."""
    NDIM = 3
    CHUNK = 1 << 16


class quadcubic(_Interp):
    """Oracle mirror of `quadcubic` (ARBInterp.py:627-1381)."""
    NDIM = 4
    CHUNK = 1 << 13


# --------------------------------------------------------------------------
# Boundary extension (no reference counterpart: README.md:35 notes that edge
# queries return NaN and plans Neumann or Dirichlet boundary conditions)
# --------------------------------------------------------------------------

def pad_nodes(values: np.ndarray, n, ndim: int, bc: str) -> np.ndarray:
    """One ghost node per face of the node grid `values` (canonical order, x
    fastest, axis lengths n), axis after axis (x, y, z[, t]):
      neumann   v(-1) = v(1),            v(n) = v(n-2)
      dirichlet v(-1) = 2 v(0) - v(1),   v(n) = 2 v(n-1) - v(n-2)
    Returns the extended grid flattened in canonical order."""
    g = values.reshape(tuple(reversed(list(n))))
    for a in range(ndim):
        g = np.moveaxis(g, ndim - 1 - a, 0)
        if bc == "neumann":
            lo, hi = g[1], g[-2]
        elif bc == "dirichlet":
            lo, hi = 2.0 * g[0] - g[1], 2.0 * g[-1] - g[-2]
        else:
            raise ValueError("unknown boundary rule %r" % (bc,))
        g = np.concatenate([lo[None], g, hi[None]], axis=0)
        g = np.moveaxis(g, 0, ndim - 1 - a)
    return np.ascontiguousarray(g).reshape(-1)


def extend_boundary(obj, bc: str):
    """Extend an oracle interpolator `obj` (fresh, no cell filled) to the whole
    sampled box: every interpolated column gets one ghost node per face
    (pad_nodes), and the grid constants become those of the extended grid
    with the original spacing: n + 2 nodes and n - 1 cells per axis, IntMin /
    IntMax = the first / last node.  The reference algorithm then runs
    unchanged on the extended grid."""
    p = obj.p
    d = obj.NDIM
    n_old = list(p.n)
    for k in list(obj.interps):
        obj.interps[k] = pad_nodes(np.asarray(obj.interps[k], dtype=np.float64), n_old, d, bc)
    n = [m + 2 for m in n_old]
    p.n = n
    p.npos = [m - 3 for m in n]
    p.stride = [int(np.prod(n[:a])) for a in range(d)]
    p.base = base_point_inds(n, d)
    p.nc = len(p.base)
    p.int_min = [p.field[0, a] for a in range(d)]
    p.int_max = [p.field[-1, a] for a in range(d)]
    obj.nc = p.nc
    obj.basePointInds = p.base
    obj._init_tables()
    return obj


# --------------------------------------------------------------------------
# Error metric used by every parity statement
# --------------------------------------------------------------------------

COND_FLOOR = 1e-4


def rel_err(out, ref, scale, floor=COND_FLOOR):
    """Relative error with a conditioning floor, elementwise:

        |out - ref| / max(|ref|, floor * S)

    where S is `condition_scale` (the magnitude of the terms the interpolant
    sums).  Outputs that cancel by more than -log10(floor) digits are judged
    against the term magnitude instead of their own tiny value.  Rows where the
    reference is NaN must be NaN in `out` too (checked separately); they are
    excluded here.  Returns (max, rms) over all finite reference entries.
    """
    out = np.asarray(out, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    scale = np.asarray(scale, dtype=np.float64)
    m = np.isfinite(ref) & np.isfinite(scale)
    if not m.any():
        return 0.0, 0.0
    den = np.maximum(np.abs(ref[m]), floor * scale[m])
    den = np.where(den > 0, den, 1.0)
    e = np.abs(out[m] - ref[m]) / den
    e = np.where(np.isnan(e), np.inf, e)
    return float(e.max()), float(np.sqrt(np.mean(e * e)))


def rel_err_normwise(out, ref, scale, floor=COND_FLOOR):
    """Normwise relative error per query with a conditioning floor:

        ||out_q - ref_q||_2 / max(||ref_q||_2, floor * ||S_q||_2)

    over the components of one output kind (the 3 field components, the
    scalar value, or the d gradient components).  Used for reduced-precision
    paths, whose error is a fraction of the output vector's size rather than
    of each component (a component crossing zero has no useful relative
    error).  Returns (max, rms) over queries with a finite reference."""
    out = np.asarray(out, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    scale = np.asarray(scale, dtype=np.float64)
    if out.ndim == 1:
        out, ref, scale = out[:, None], ref[:, None], scale[:, None]
    m = np.isfinite(ref).all(axis=1) & np.isfinite(scale).all(axis=1)
    if not m.any():
        return 0.0, 0.0
    num = np.linalg.norm(out[m] - ref[m], axis=1)
    den = np.maximum(np.linalg.norm(ref[m], axis=1), floor * np.linalg.norm(scale[m], axis=1))
    den = np.where(den > 0, den, 1.0)
    e = num / den
    e = np.where(np.isnan(e), np.inf, e)
    return float(e.max()), float(np.sqrt(np.mean(e * e)))

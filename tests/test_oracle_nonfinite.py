"""Non-finite field nodes in the oracle, pinned by the closed form.

The reference forms a cell's coefficients as alpha = A . F[neighbourInd]
(ARBInterp.py:599-602, 4D :1351-1354).  A has a zero in every column and
0 * NaN = 0 * Inf = NaN, so a NaN or +-Inf node value makes every output of
every cell whose 4^d stencil holds that node NaN, and changes no other cell.
The expected mask is formed here by brute force over cell and node indices
(independently of the oracle's neighbour offsets), and the outputs of the
untouched cells must be bitwise those of the same field without the bad node.
"""
import itertools

import numpy as np
import pytest

from arbinterp_amd import workloads as W
from oracle import arbinterp_oracle as O

# the oracle's matmuls warn about the NaN they produce on purpose here
pytestmark = pytest.mark.filterwarnings("ignore::RuntimeWarning")


def _field(d, n, ncomp, seed):
    rng = np.random.default_rng(seed)
    h = [0.1, 0.13, 0.09, 0.2][:d]
    axes = W.grid_axes(n, h)
    mesh = np.meshgrid(*axes, indexing="ij")
    vals = []
    for c in range(ncomp):
        v = np.zeros(n)
        for k in range(3):
            ctr = rng.uniform(0, 1, size=d) * np.array([a[-1] for a in axes])
            r2 = sum((mesh[a] - ctr[a]) ** 2 for a in range(d))
            v += rng.uniform(0.5, 2.0) * np.exp(-r2 / 0.2) * (1 if c % 2 == 0 else -1)
        vals.append(v)
    return axes, vals


def _cell_points(axes, d, per_cell, rng):
    """per_cell random points inside every interpolation cell, and their cell."""
    n = [len(a) for a in axes]
    npos = [k - 3 for k in n]
    pts, cells = [], []
    for c in itertools.product(*[range(k) for k in npos]):
        for _ in range(per_cell):
            u = rng.uniform(0.02, 0.98, size=d)
            pts.append([axes[a][c[a] + 1] + u[a] * (axes[a][c[a] + 2] - axes[a][c[a] + 1])
                        for a in range(d)])
            cells.append(c)
    return np.array(pts), np.array(cells)


def _expected_mask(cells, node):
    """Cells whose stencil (nodes c .. c + 3 per axis, cell c's base node is
    c + 1) holds `node`."""
    node = np.asarray(node)
    return np.all((node[None, :] >= cells) & (node[None, :] <= cells + 3), axis=1)


CASES = [
    (3, (9, 8, 7), None, (4, 3, 2), np.nan),
    (3, (9, 8, 7), None, (0, 0, 0), np.inf),
    (3, (9, 8, 7), "both", (5, 4, 6), -np.inf),
    (3, (9, 8, 7), "vector", (3, 7, 3), np.nan),
    (3, (9, 8, 7), "norm", (8, 2, 4), np.nan),
    (4, (6, 5, 6, 5), None, (2, 3, 1, 2), np.nan),
    (4, (6, 5, 6, 5), "both", (1, 1, 4, 3), np.inf),
]


@pytest.mark.parametrize("d,n,mode,node,bad", CASES)
def test_nonfinite_node_poisons_exactly_its_stencil_cells(d, n, mode, node, bad):
    ncomp = 1 if mode is None else 3
    axes, vals = _field(d, n, ncomp, seed=sum(n) + len(str(mode)))
    rows_ok = W.grid_rows(axes, vals)
    vals_bad = [v.copy() for v in vals]
    comp = 0 if ncomp == 1 else (1 if mode == "both" else 2)
    vals_bad[comp][node] = bad
    rows_bad = W.grid_rows(axes, vals_bad, np.random.default_rng(5))
    kw = {} if mode is None else {"mode": mode}
    cls = O.tricubic if d == 3 else O.quadcubic
    ok = cls(rows_ok.copy(), "quiet", **kw)
    bd = cls(rows_bad.copy(), "quiet", **kw)
    q, cells = _cell_points(axes, d, 2, np.random.default_rng(6))
    want = _expected_mask(cells, node)
    assert want.any() and not want.all()
    r_ok = ok.Query(q.copy())
    r_bd = bd.Query(q.copy())
    r_ok = r_ok if isinstance(r_ok, tuple) else (r_ok,)
    r_bd = r_bd if isinstance(r_bd, tuple) else (r_bd,)
    for k, (a, b) in enumerate(zip(r_ok, r_bd)):
        a = a.reshape(len(q), -1)
        b = b.reshape(len(q), -1)
        assert not np.isnan(a).any()
        if mode in ("vector", "both") and k == 0:
            # components: only the poisoned component's column is NaN
            got = np.isnan(b)
            assert np.array_equal(got[:, comp], want), (mode, k)
            others = [c for c in range(3) if c != comp]
            assert not got[:, others].any()
            assert np.array_equal(a[:, others], b[:, others])
            assert np.array_equal(a[~want], b[~want])
        else:
            got = np.isnan(b).any(axis=1)
            assert np.array_equal(got, want), (mode, k)
            assert np.isnan(b[want]).all()
            assert np.array_equal(a[~want], b[~want])

"""The oracle's boundary extension (oracle.extend_boundary / pad_nodes), pinned
by what the mathematics fixes rather than by its own formula:

  * a field even about every face (products of cos) has exact Neumann
    ghosts, and one odd about every face (products of sin) has exact
    Dirichlet ghosts; then the edge cells are as accurate as the interior
    ones, while the other rule leaves an O(h) slope error at the edges;
  * Neumann: the normal slope at the edge node is zero;
  * Dirichlet: a linear field is reproduced on the whole box;
  * inside the reference's box the extension changes nothing;
  * the grid constants describe the extended grid.
"""
import numpy as np
import pytest

from arbinterp_amd import workloads as W
from oracle import arbinterp_oracle as O


def _field(kind, n=(14, 12, 10), h=(0.1, 0.12, 0.15), origin=(0.3, -0.2, 0.05), nvec=0):
    axes = W.grid_axes(n, h, origin)
    L = [axes[a][-1] - axes[a][0] for a in range(3)]
    X = np.meshgrid(*[(axes[a] - axes[a][0]) / L[a] for a in range(3)], indexing="ij")
    if kind == "even":
        v = np.cos(np.pi * X[0]) * np.cos(np.pi * X[1]) * np.cos(np.pi * X[2])
    elif kind == "odd":
        v = np.sin(np.pi * X[0]) * np.sin(np.pi * X[1]) * np.sin(np.pi * X[2])
    else:                                   # linear
        v = 0.7 + 1.3 * X[0] - 0.4 * X[1] + 2.1 * X[2]
    vals = [v] if nvec == 0 else [v * (1 + 0.1 * j) for j in range(3)]
    return W.grid_rows(axes, vals, np.random.default_rng(3)), axes


def _make(rows, bc, **kw):
    o = O.tricubic(rows.copy(), "quiet", **kw)
    if bc:
        O.extend_boundary(o, bc)
    return o


def _box_queries(o, rng, m):
    lo = np.array(o.p.int_min, dtype=np.float64)
    hi = np.array(o.p.int_max, dtype=np.float64)
    # strictly below IntMax: a query at IntMax aliases into the next row (as
    # in the reference), which is not what these tests look at
    return lo + (hi - lo) * rng.uniform(0.0, 1.0 - 1e-9, size=(m, 3))


def _truth(kind, axes, q):
    L = [axes[a][-1] - axes[a][0] for a in range(3)]
    X = [(q[:, a] - axes[a][0]) / L[a] for a in range(3)]
    if kind == "even":
        return np.cos(np.pi * X[0]) * np.cos(np.pi * X[1]) * np.cos(np.pi * X[2])
    if kind == "odd":
        return np.sin(np.pi * X[0]) * np.sin(np.pi * X[1]) * np.sin(np.pi * X[2])
    return 0.7 + 1.3 * X[0] - 0.4 * X[1] + 2.1 * X[2]


def _edge_mask(o, q, axes):
    """Rows in a cell that uses a ghost node (outside the reference's box)."""
    m = np.zeros(len(q), dtype=bool)
    for a in range(3):
        m |= (q[:, a] < axes[a][1]) | (q[:, a] > axes[a][-2])
    return m


@pytest.mark.parametrize("kind,bc,other", [("even", "neumann", "dirichlet"),
                                           ("odd", "dirichlet", "neumann")])
def test_exact_ghosts_keep_edge_accuracy(kind, bc, other):
    rows, axes = _field(kind)
    rng = np.random.default_rng(11)
    o = _make(rows, bc)
    q = _box_queries(o, rng, 20000)
    edge = _edge_mask(o, q, axes)
    assert edge.sum() > 2000 and (~edge).sum() > 2000
    err = np.abs(o.Query(q.copy())[0][:, 0] - _truth(kind, axes, q))
    e_in, e_edge = err[~edge].max(), err[edge].max()
    assert e_edge <= 3.0 * e_in, (e_edge, e_in)
    # the other rule has the wrong slope at the edge node: O(h) errors there
    w = _make(rows, other)
    err_w = np.abs(w.Query(q.copy())[0][:, 0] - _truth(kind, axes, q))
    assert err_w[edge].max() > 10.0 * e_in, (err_w[edge].max(), e_in)


def test_neumann_zero_normal_slope():
    rows, axes = _field("linear")            # nonzero slope everywhere inside
    o = _make(rows, "neumann")
    rng = np.random.default_rng(5)
    q = _box_queries(o, rng, 500)
    q[:, 0] = axes[0][0]                     # on the x = first node face
    g = o.Query(q.copy())[1]
    sc = o.condition_scale(q)[1]
    assert np.all(np.abs(g[:, 0]) <= 1e-13 * sc[:, 0])
    # the same field without the extension has slope 1.3 / L_x there
    inner = q.copy()
    inner[:, 0] = axes[0][1]
    gi = o.Query(inner.copy())[1]
    assert np.all(np.abs(gi[:, 0]) > 0.5)


def test_dirichlet_reproduces_linear_fields():
    rows, axes = _field("linear")
    o = _make(rows, "dirichlet")
    rng = np.random.default_rng(6)
    q = _box_queries(o, rng, 5000)
    q[:100, 0] = axes[0][0]
    v, g = o.Query(q.copy())
    sc = o.condition_scale(q)
    L = [axes[a][-1] - axes[a][0] for a in range(3)]
    assert np.all(np.abs(v[:, 0] - _truth("linear", axes, q)) <= 1e-12 * sc[0][:, 0])
    slope = np.array([1.3 / L[0], -0.4 / L[1], 2.1 / L[2]])
    assert np.all(np.abs(g - slope[None, :]) <= 1e-12 * sc[1])


@pytest.mark.parametrize("bc", ["neumann", "dirichlet"])
@pytest.mark.parametrize("mode", [None, "both"])
def test_interior_unchanged(bc, mode):
    kw = {} if mode is None else {"mode": mode}
    rows, axes = _field("even", nvec=3 if mode else 0)
    base = _make(rows, None, **kw)
    ext = _make(rows, bc, **kw)
    rng = np.random.default_rng(7)
    lo = np.array(base.p.int_min)
    hi = np.array(base.p.int_max)
    q = lo + (hi - lo) * rng.uniform(0.0, 1.0 - 1e-9, size=(3000, 3))
    a = base.Query(q.copy())
    b = ext.Query(q.copy())
    sc = base.condition_scale(q)
    sc = sc if isinstance(sc, tuple) else (sc,)
    for x, y, s in zip(a, b, sc):
        assert np.all(np.abs(x - y) <= 1e-14 * s.reshape(x.shape))


def test_extended_grid_constants():
    rows, axes = _field("even")
    o = _make(rows, "dirichlet")
    n = [len(ax) for ax in axes]
    assert o.p.npos == [m - 1 for m in n]
    assert o.nc == int(np.prod([m - 1 for m in n]))
    assert o.p.int_min == [ax[0] for ax in axes] and o.p.int_max == [ax[-1] for ax in axes]
    # the ghost layer: brute force over the extended grid for a tiny field
    v = np.random.default_rng(9).normal(size=4 * 5 * 6)
    g = O.pad_nodes(v, [4, 5, 6], 3, "neumann").reshape(8, 7, 6)        # (z, y, x)
    core = v.reshape(6, 5, 4)
    def src(i, m):                   # extended index -> original index (even rule)
        return 1 if i == 0 else (m - 2 if i == m + 1 else i - 1)
    for z in range(8):
        for y in range(7):
            for x in range(6):
                assert g[z, y, x] == core[src(z, 6), src(y, 5), src(x, 4)]

"""Coefficients formed on the fly when the table does not fit.

The reference fills only the cells its queries touch (ARBInterp.py:376-377,
:489-490).  Here a field whose coefficient table does not fit in device
memory, or exceeds a size limit (table_limit / arb_set_table_limit, which
forces the path on small grids), answers every query from fused bucketed
kernels that form each bucket's coefficients in LDS with the table build's
arithmetic.  Checked: the outputs, cell indices and NaN rows are bitwise
those of the same field with its table, for every mode, 3D and 4D, FP64 and
mixed, on the bucketed and direct paths (both become the on-the-fly
bucketed path) and for single points, also with a NaN node; a limit set
after the table was built releases it; and the calls that need the table
itself raise MemoryError.
"""
import numpy as np
import pytest

from arbinterp_amd import _native as N
from arbinterp_amd import workloads as W

pytestmark = [pytest.mark.gpu, pytest.mark.filterwarnings("ignore::RuntimeWarning")]


@pytest.fixture(scope="module")
def G():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a ROCm GPU"
    from arbinterp_amd import ARBInterp as G
    return G


def _field(d, n, ncomp, seed, nan_node=None):
    rng = np.random.default_rng(seed)
    h = [0.1, 0.12, 0.09, 0.15][:d]
    axes = W.grid_axes(n, h)
    mesh = np.meshgrid(*axes, indexing="ij")
    ext = np.array([a[-1] for a in axes])
    vals = []
    for c in range(ncomp):
        v = np.zeros(n)
        for _ in range(3):
            ctr = rng.uniform(0.2, 0.8, size=d) * ext
            r2 = sum(((mesh[a] - ctr[a]) / (0.4 * ext[a])) ** 2 for a in range(d))
            v += rng.uniform(0.5, 2.0) * np.exp(-r2) * (1.0 if c != 1 else -1.0)
        vals.append(v + (0.5 if c == 0 else 0.0))
    if nan_node is not None:
        vals[0][nan_node] = np.nan
    return W.grid_rows(axes, vals, rng), axes


def _queries(axes, nq, seed):
    rng = np.random.default_rng(seed)
    lo = np.array([a[1] for a in axes])
    hi = np.array([a[-2] for a in axes])
    h = np.array([a[1] - a[0] for a in axes])
    q = rng.uniform(lo - 0.5 * h, hi + 0.5 * h, size=(nq, len(axes)))
    q[:3] = np.nan
    return q


def _run(t, q, flags):
    import torch
    qd = torch.from_numpy(np.ascontiguousarray(q)).cuda()
    out = t.query_device(qd, nan_rows=True, out_f64=True, cells=True, flags=flags)
    torch.cuda.synchronize()
    return [x.cpu().numpy() if x is not None else None for x in out] + [qd.cpu().numpy()]


def _bitwise(a, b, tag):
    for k, (x, y) in enumerate(zip(a, b)):
        if x is None:
            assert y is None, tag
            continue
        assert x.shape == y.shape, (tag, k)
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), (tag, k)


CASES = [
    (3, "fp64", "both"), (3, "fp64", "scalar"), (3, "fp64", "vector"), (3, "fp64", "norm"),
    (3, "mixed", "both"), (4, "fp64", "scalar"), (4, "mixed", "scalar"), (4, "fp64", "both"),
    (4, "mixed", "both"),
]


@pytest.mark.parametrize("d,prec,mode", CASES)
def test_on_the_fly_equals_table(G, d, prec, mode):
    n = (16, 14, 13) if d == 3 else (9, 8, 8, 7)
    ncomp = 1 if mode == "scalar" else 3
    rows, axes = _field(d, n, ncomp, seed=d * 10 + len(mode))
    kw = {} if mode == "scalar" else {"mode": mode}
    cls = G.tricubic if d == 3 else G.quadcubic
    tag = "%dD %s %s" % (d, prec, mode)
    print(tag, flush=True)
    t_tab = cls(rows.copy(), "quiet", precision=prec, **kw)
    t_otf = cls(rows.copy(), "quiet", precision=prec, table_limit=1, **kw)
    q = _queries(axes, 20_000, seed=5)
    # mixed 3D fields read node tiles by default (no table either way); the
    # table family is compared through ARB_Q_TABLE
    base = N.ARB_Q_TABLE if (d == 3 and prec == "mixed") else 0
    for name, fl in (("bucketed", N.ARB_Q_SORTED), ("direct", N.ARB_Q_DIRECT)):
        ref = _run(t_tab, q, base | fl)
        got = _run(t_otf, q, base | fl)
        _bitwise(got, ref, tag + " " + name)
    info = t_otf.field_info()
    assert info.on_the_fly == 1 and info.table_built == 0, tag
    t_otf.allCoeffs()                               # not built: a no-op, no error
    assert t_otf.field_info().table_built == 0
    if not (d == 3 and prec == "mixed"):
        # single points inside the box: one-row batches on the on-the-fly
        # path, the table path's bits
        lo = np.array([a[1] for a in axes])
        hi = np.array([a[-2] for a in axes])
        pts = lo + (hi - lo) * np.random.default_rng(9).uniform(0.05, 0.95, size=(3, d))
        for r in range(3):
            a = t_otf.Query(pts[r].copy())
            b = t_tab.Query(pts[r].copy())
            a = a if isinstance(a, tuple) else (a,)
            b = b if isinstance(b, tuple) else (b,)
            for x, y in zip(a, b):
                assert np.array_equal(np.atleast_1d(x).view(np.uint8),
                                      np.atleast_1d(y).view(np.uint8)), (tag, r)
        with pytest.raises(MemoryError):
            t_otf.coefficient_table(0, 4)
    t_tab.close()
    t_otf.close()


@pytest.mark.parametrize("d", [3, 4])
def test_on_the_fly_nan_node_and_late_limit(G, d):
    """A NaN node poisons the same cells on the fly as in the table; a limit
    set after the table was built releases it, and later queries still give
    the table's bits."""
    n = (14, 13, 12) if d == 3 else (8, 8, 7, 7)
    node = (6, 5, 4) if d == 3 else (4, 3, 3, 3)
    rows, axes = _field(d, n, 1, seed=7 + d, nan_node=node)
    cls = G.tricubic if d == 3 else G.quadcubic
    t = cls(rows.copy(), "quiet")
    q = _queries(axes, 15_000, seed=6)
    ref = _run(t, q, N.ARB_Q_SORTED)
    assert np.isnan(ref[1]).sum() > 3               # poisoned cells were queried
    assert t.field_info().table_built == 1
    t.set_table_limit(1)
    info = t.field_info()
    assert info.on_the_fly == 1 and info.table_built == 0
    got = _run(t, q, N.ARB_Q_SORTED)
    _bitwise(got, ref, "%dD late limit" % d)
    t.set_table_limit(0)                            # lifted: the table is built again
    again = _run(t, q, N.ARB_Q_DIRECT)
    _bitwise(again, ref, "%dD limit lifted" % d)
    assert t.field_info().on_the_fly == 0 and t.field_info().table_built == 1
    t.close()


def test_on_the_fly_refuses_table_readers(G):
    import torch
    from arbinterp_amd import autograd as AG
    from arbinterp_amd import batch as B
    rows, axes = _field(3, (12, 11, 10), 1, seed=3)
    t = G.tricubic(rows.copy(), "quiet", table_limit=1)
    q = torch.from_numpy(_queries(axes, 1000, seed=2)[3:]).cuda()
    fld = AG.DifferentiableField(t)
    U, Gr = fld(q.clone().requires_grad_(True))     # the forward works on the fly
    with pytest.raises(MemoryError):
        (U.sum() + Gr.sum()).backward()
    with pytest.raises(MemoryError):
        B.FieldBatch([t])
    x = q[:10].clone()
    with pytest.raises(MemoryError):
        B.push_rk4(t, x, torch.zeros_like(x), dt=1e-3, nsteps=2)
    torch.cuda.synchronize()
    t.close()


@pytest.mark.parametrize("d,bc", [(3, "neumann"), (3, "dirichlet"), (4, "neumann")])
def test_on_the_fly_boundary_fields(G, d, bc):
    """Boundary-extended fields (one ghost layer per face) without a table:
    the fused kernels read the extended node grid, so the whole sampled box
    gives the table's bits."""
    n = (12, 11, 10) if d == 3 else (8, 7, 7, 6)
    rows, axes = _field(d, n, 1, seed=20 + d)
    cls = G.tricubic if d == 3 else G.quadcubic
    tag = "%dD %s" % (d, bc)
    print(tag, flush=True)
    t_tab = cls(rows.copy(), "quiet", boundary=bc)
    t_otf = cls(rows.copy(), "quiet", boundary=bc, table_limit=1)
    # positions over the whole sampled box (the rim the boundary makes valid)
    rng = np.random.default_rng(4)
    lo = np.array([a[0] for a in axes])
    hi = np.array([a[-1] for a in axes])
    q = rng.uniform(lo, hi, size=(12_000, d))
    for name, fl in (("bucketed", N.ARB_Q_SORTED), ("direct", N.ARB_Q_DIRECT)):
        _bitwise(_run(t_otf, q, fl), _run(t_tab, q, fl), tag + " " + name)
    got = _run(t_otf, q, N.ARB_Q_SORTED)
    assert np.isfinite(got[1]).all(), tag                 # no NaN rim
    assert t_otf.field_info().on_the_fly == 1
    t_tab.close()
    t_otf.close()

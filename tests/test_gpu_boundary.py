"""GPU parity of the boundary extension (boundary='neumann' | 'dirichlet')
against the oracle's restatement (oracle.extend_boundary, pinned by
tests/test_oracle_boundary.py), plus the properties the GPU path must keep:
bitwise unchanged results inside the reference's box, an exactly zero normal
slope on a Neumann face, the extended grid's attributes, and loud errors for
what a boundary-extended field does not offer."""
import numpy as np
import pytest

from arbinterp_amd import workloads as W
from oracle import arbinterp_oracle as O

pytestmark = pytest.mark.gpu
TOL64 = 1e-12
# Mixed precision: in the last cells before a Neumann face the zero-slope rule
# bends the cubic (large u^2, u^3 coefficients), and the FP32 gradient near
# u = 1 cancels those terms; measured 2.9e-6 normwise (floor 1e-4 S), which is
# 9e-10 of the condition scale S.  Both bounds are checked: normwise with the
# usual floor, and relative to S itself (a few FP32 roundings).
TOLMIX = 5e-6
TOLMIX_S = 2e-7


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a ROCm GPU"
    from arbinterp_amd import ARBInterp as G
    return G


def _rows3(nvec, seed=4, n=(14, 12, 10), h=(0.1, 0.12, 0.15), origin=(0.3, -0.2, 0.05)):
    rng = np.random.default_rng(seed)
    axes = W.grid_axes(n, h, origin)
    X = np.meshgrid(*axes, indexing="ij")
    base = np.sin(1.7 * X[0] + 0.3) * np.cos(1.1 * X[1] - 0.2) * np.exp(0.4 * X[2])
    vals = [base] if nvec == 0 else [base * (1 + 0.2 * j) + 0.1 * j * X[j] for j in range(3)]
    return W.grid_rows(axes, vals, rng), axes


def _queries(lo, hi, rng, m, axes=None):
    d = len(lo)
    q = lo + (hi - lo) * rng.uniform(0.0, 1.0 - 1e-9, size=(m, d))
    q[:64, 0] = lo[0]                               # the x = first node face
    q[64:96, :] = lo                                # the lower corner
    return q


def _check(got, ref, sc, precision, label):
    got = got if isinstance(got, tuple) else (got,)
    ref = ref if isinstance(ref, tuple) else (ref,)
    sc = sc if isinstance(sc, tuple) else (sc,)
    assert len(got) == len(ref), label
    for i, (g, r, s) in enumerate(zip(got, ref, sc)):
        g = np.asarray(g, dtype=np.float64)
        s = np.asarray(s, dtype=np.float64).reshape(r.shape)
        assert g.shape == r.shape, (label, i)
        assert np.array_equal(np.isnan(g), np.isnan(r)), (label, i)
        ok = ~np.isnan(r)
        if precision == "fp64":
            mx, _ = O.rel_err(g[ok], r[ok], s[ok])
            assert mx <= TOL64, (label, i, mx)
        else:
            r2, g2, s2 = (x.reshape(len(x), -1) for x in (r, g, s))
            rows = ok.reshape(len(ok), -1).all(axis=1)
            mx, _ = O.rel_err_normwise(g2[rows], r2[rows], s2[rows])
            assert mx <= TOLMIX, (label, i, mx)
            ms, _ = O.rel_err_normwise(g2[rows], r2[rows], s2[rows], floor=1.0)
            assert ms <= TOLMIX_S, (label, i, ms)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("bc", ["neumann", "dirichlet"])
@pytest.mark.parametrize("mode", [None, "both", "norm", "vector"])
def test_boundary_3d_matches_oracle(dev, bc, mode, precision):
    rows, axes = _rows3(0 if mode is None else 3)
    kw = {} if mode is None else {"mode": mode}
    t = dev.tricubic(rows.copy(), "quiet", boundary=bc, precision=precision, **kw)
    o = O.extend_boundary(O.tricubic(rows.copy(), "quiet", **kw), bc)
    lo, hi = np.array(o.p.int_min), np.array(o.p.int_max)
    assert np.array_equal(t._lo, lo) and np.array_equal(t._hi, hi)
    q = _queries(lo, hi, np.random.default_rng(21), 4000)
    got = t.Query(q.copy())
    ref = o.Query(q.copy())
    _check(got, ref, o.condition_scale(q), precision, (bc, mode, precision))
    assert np.array_equal(t.queryInds, o.queryInds)
    t.close()


@pytest.mark.parametrize("bc", ["neumann", "dirichlet"])
def test_boundary_bucketed_path(dev, bc):
    """Mixed 'both' batches on the bucketed path.  Dirichlet-extended fields
    are evaluated from node tiles of the extended grid (ghost layers
    included).  Neumann-extended fields keep the coefficient table: its
    FP64-formed slopes make the normal slope on the faces exactly zero (node
    tiles would leave 2^-24 of the tile differences there, about 1e-7 at a
    corner whose reference gradient is exactly zero)."""
    import torch
    from arbinterp_amd import _native as NN
    rows, axes = _rows3(3)
    t = dev.tricubic(rows.copy(), "quiet", mode="both", boundary=bc, precision="mixed")
    o = O.extend_boundary(O.tricubic(rows.copy(), "quiet", mode="both"), bc)
    lo, hi = np.array(o.p.int_min), np.array(o.p.int_max)
    q = _queries(lo, hi, np.random.default_rng(27), 80_000 if bc == "dirichlet" else 4000)
    qd = torch.from_numpy(q).cuda()
    c, nv, g, cells = t.query_device(qd, out_f64=True, cells=True, flags=NN.ARB_Q_SORTED)
    torch.cuda.synchronize()
    assert (t.field_info().tiles_bytes > 0) == (bc == "dirichlet")
    got = (c.cpu().numpy(), nv.cpu().numpy()[:, None], g.cpu().numpy())
    ref = o.Query(q.copy())
    _check(got, ref, o.condition_scale(q), "mixed", (bc, "bucketed"))
    assert np.array_equal(cells.cpu().numpy(), o.queryInds)
    if bc == "neumann":
        qf = q.copy()
        qf[:, 0] = lo[0]
        _, _, gf = t.query_device(torch.from_numpy(qf).cuda(), flags=NN.ARB_Q_SORTED)
        torch.cuda.synchronize()
        assert torch.all(gf[:, 0] == 0.0)
    t.close()


@pytest.mark.parametrize("bc", ["neumann", "dirichlet"])
def test_boundary_4d_matches_oracle(dev, golden, bc):
    fld = golden("vector4d_modes")["field"]
    for kw in ({"mode": "both"}, {}):
        t = dev.quadcubic(fld.copy(), "quiet", boundary=bc, **kw)
        o = O.extend_boundary(O.quadcubic(fld.copy(), "quiet", **kw), bc)
        lo, hi = np.array(o.p.int_min), np.array(o.p.int_max)
        q = _queries(lo, hi, np.random.default_rng(22), 2000)
        _check(t.Query(q.copy()), o.Query(q.copy()), o.condition_scale(q), "fp64", (bc, kw))
        assert np.array_equal(t.queryInds, o.queryInds)
        t.close()


@pytest.mark.parametrize("bc", ["neumann", "dirichlet"])
def test_boundary_interior_unchanged(dev, bc):
    """Inside the reference's box every cell's 64 nodes are original nodes, so
    the extended field's coefficients of those cells are the plain field's,
    bit for bit (cell (i, j, k) becomes (i+1, j+1, k+1)).  Query results there
    differ only by the rounding of the local coordinates, (q - x0)/h instead
    of (q - x1)/h."""
    rows, axes = _rows3(3)
    n = [len(ax) for ax in axes]
    o = O.tricubic(rows.copy(), "quiet", mode="both")
    for precision in ("fp64", "mixed"):
        a = dev.tricubic(rows.copy(), "quiet", mode="both", precision=precision)
        b = dev.tricubic(rows.copy(), "quiet", mode="both", precision=precision, boundary=bc)
        ta = a.coefficient_table()[:-1].reshape(n[2] - 3, n[1] - 3, n[0] - 3, 4, 64)
        tb = b.coefficient_table()[:-1].reshape(n[2] - 1, n[1] - 1, n[0] - 1, 4, 64)
        assert np.array_equal(ta, tb[1:-1, 1:-1, 1:-1]), (bc, precision)
        q = _queries(a._lo, a._hi, np.random.default_rng(23), 3000)
        _check(b.Query(q.copy()), a.Query(q.copy()), o.condition_scale(q), precision,
               (bc, precision))
        a.close()
        b.close()


def test_neumann_normal_slope_exactly_zero(dev):
    rows, axes = _rows3(0)
    t = dev.tricubic(rows.copy(), "quiet", boundary="neumann")
    q = _queries(t._lo, t._hi, np.random.default_rng(25), 500)
    q[:, 0] = t._lo[0]
    _, g = t.Query(q.copy())
    assert np.all(g[:, 0] == 0.0)
    t.close()


def test_boundary_attributes_and_errors(dev):
    rows, axes = _rows3(3)
    n = [len(ax) for ax in axes]
    t = dev.tricubic(rows.copy(), "quiet", mode="both", boundary="dirichlet")
    assert t.boundary == "dirichlet"
    assert list(t.nPos) == [m - 1 for m in n]
    assert t.nc == int(np.prod([m - 1 for m in n]))
    assert t.xIntMin == axes[0][0] and t.zIntMax == axes[2][-1] and t.hy == t._hh[1]
    assert t._info.boundary == 2
    # a single point on the first node corner is inside the extended box
    pt = np.array([axes[0][0], axes[1][0], axes[2][0]])
    comps, nv, gv = t.Query(pt.copy())
    assert np.all(np.isfinite(comps)) and np.isfinite(nv) and t.queryInd == 0
    from arbinterp_amd import autograd as AG
    with pytest.raises(ValueError):
        AG.DifferentiableField(t)
    with pytest.raises(ValueError):
        dev.tricubic(rows.copy(), "quiet", boundary="periodic")
    t.close()

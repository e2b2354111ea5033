"""Non-finite field nodes on the GPU: the reference's NaN mask, exactly.

In the reference a NaN or +-Inf node value makes every output of every cell
whose 4^d stencil holds the node NaN (alpha = A . F[neighbourInd],
ARBInterp.py:599-602 / :1351-1354; A has a zero in every column), and leaves
every other cell as it was; the oracle does the same (pinned by the closed
form in test_oracle_nonfinite.py).  Checked here against the oracle on every
query path (bucketed and direct, the FP64 table, the fused kernel, the mixed
node tiles and the mixed FP32 table, single points), in every mode, 3D and
4D, with the bad node at a bucket's model centre (the tile's middle node),
one and two nodes from it, at the tile's first and last node and at a grid
corner: identical cell indices and NaN masks per output column, and the
usual bounds on every finite output (FP64 1e-12 elementwise, mixed 1e-6
normwise, both with the oracle's conditioning floor).
"""
import numpy as np
import pytest

from arbinterp_amd import _native as N
from arbinterp_amd import workloads as W
from oracle import arbinterp_oracle as O

pytestmark = [pytest.mark.gpu, pytest.mark.filterwarnings("ignore::RuntimeWarning")]
TOL64 = 1e-12
TOLMIX = 1e-6


@pytest.fixture(scope="module")
def G():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a ROCm GPU"
    from arbinterp_amd import ARBInterp as G
    return G


def _field(d, n, ncomp, seed):
    """Sum of Gaussians per component on a grid with spacings h."""
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
            w = rng.uniform(0.3, 0.6) * ext.mean()
            r2 = sum((mesh[a] - ctr[a]) ** 2 for a in range(d))
            v += rng.uniform(0.5, 2.0) * np.exp(-r2 / (2 * w * w)) * (1.0 if c != 1 else -1.0)
        vals.append(v + (1.5 if c == 0 else 0.0))
    return axes, vals


def _queries(axes, node, nq, rng):
    """Points within 4 cells of `node` (clipped to the interior) and over the
    whole interior, plus a few outside the box."""
    d = len(axes)
    lo = np.array([a[1] for a in axes])
    hi = np.array([a[-2] for a in axes])
    h = np.array([a[1] - a[0] for a in axes])
    ctr = np.array([axes[a][node[a]] for a in range(d)])
    near = rng.uniform(np.maximum(lo, ctr - 4 * h), np.minimum(hi, ctr + 4 * h), size=(nq, d))
    far = rng.uniform(lo, hi, size=(nq // 4, d))
    out = rng.uniform(lo - h, hi + h, size=(64, d))
    return np.concatenate([near, far, out])


def _compare(got, ref, scale, cells_g, cells_o, prec, tag):
    assert np.array_equal(cells_g, cells_o), tag + ": cell indices"
    for k, (a, r, s) in enumerate(zip(got, ref, scale)):
        a = np.asarray(a, dtype=np.float64).reshape(len(r), -1)
        r = np.asarray(r, dtype=np.float64).reshape(len(r), -1)
        s = np.asarray(s, dtype=np.float64).reshape(len(r), -1)
        assert np.array_equal(np.isnan(a), np.isnan(r)), "%s out%d: NaN mask (%d vs %d)" % (
            tag, k, np.isnan(a).sum(), np.isnan(r).sum())
        if prec == "fp64":
            mx, _ = O.rel_err(a, r, s)
            assert mx <= TOL64, "%s out%d: rel err %.3e" % (tag, k, mx)
        else:
            fin = np.isfinite(r).all(axis=1)
            if a.shape[1] > 1 and k == 0 and not fin.all():
                # components: judge each finite component column on its own
                for c in range(a.shape[1]):
                    fc = np.isfinite(r[:, c])
                    mx, _ = O.rel_err_normwise(a[fc, c:c + 1], r[fc, c:c + 1], s[fc, c:c + 1])
                    assert mx <= TOLMIX, "%s out%d col %d: rel err %.3e" % (tag, k, c, mx)
            else:
                mx, _ = O.rel_err_normwise(a[fin], r[fin], s[fin])
                assert mx <= TOLMIX, "%s out%d: rel err %.3e" % (tag, k, mx)


def _gpu_query(t, q, flags):
    import torch
    qd = torch.from_numpy(np.ascontiguousarray(q)).cuda()
    out = t.query_device(qd, out_f64=True, cells=True, flags=flags)
    torch.cuda.synchronize()
    c, nv, g, cells = [x.cpu().numpy() if x is not None else None for x in out]
    return c, nv, g, cells


def _outputs_for(mode, c, nv, g):
    if mode == "vector":
        return [c]
    if mode in ("norm", "scalar"):
        return [nv.reshape(-1, 1), g]
    return [c, nv.reshape(-1, 1), g]


def _placements(t, G):
    """Bad-node positions relative to the model centre of the bucket at block
    coordinates (1, 1, 1): its tile starts at node c0 = block, and its middle
    node is c0 + (block + 3) // 2 per axis, kept two nodes inside the grid."""
    import torch
    info = t.field_info()
    if info.block[0] == 0:   # fix the bucket geometry with one bucketed query
        _gpu_query(t, np.array([[t.xIntMin, t.yIntMin, t.zIntMin]]), N.ARB_Q_SORTED)
        info = t.field_info()
    blk = np.array([info.block[a] for a in range(3)])
    n = np.array([info.n[a] for a in range(3)])
    c0 = blk.copy()
    m = np.clip(c0 + (blk + 3) // 2, 2, n - 3)
    tile_last = np.minimum(c0 + blk + 2, n - 1)
    return {
        "centre": tuple(m), "centre+1x": tuple(m + [1, 0, 0]), "centre-1y": tuple(m - [0, 1, 0]),
        "centre+2z": tuple(m + [0, 0, 2]), "centre-2x": tuple(m - [2, 0, 0]),
        "tile_first": tuple(c0), "tile_last": tuple(tile_last), "grid_corner": (0, 0, 0),
    }


MODES3 = ["both", "vector", "norm", "scalar"]
BAD = [np.nan, np.inf, -np.inf]


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
@pytest.mark.parametrize("mode", MODES3)
def test_nonfinite_nodes_3d(G, mode, prec):
    n = (22, 20, 18)
    ncomp = 1 if mode == "scalar" else 3
    axes, vals = _field(3, n, ncomp, seed=31 + MODES3.index(mode))
    kw = {} if mode == "scalar" else {"mode": mode}
    probe = G.tricubic(W.grid_rows(axes, vals), "quiet", precision=prec, **kw)
    places = _placements(probe, G)
    probe.close()
    names = list(places) if prec == "mixed" else ["centre", "tile_first"]
    rng = np.random.default_rng(17)
    for i, name in enumerate(names):
        node = places[name]
        bad = BAD[i % 3]
        comp = 0 if ncomp == 1 else i % 3
        vb = [v.copy() for v in vals]
        vb[comp][node] = bad
        rows = W.grid_rows(axes, vb, np.random.default_rng(i))
        tag = "%s %s %s comp%d=%s" % (prec, mode, name, comp, bad)
        print(tag, flush=True)
        t = G.tricubic(rows.copy(), "quiet", precision=prec, **kw)
        o = O.tricubic(rows.copy(), "quiet", **kw)
        q = _queries(axes, node, 3000, rng)
        ref = o.Query(q.copy())
        ref = list(ref) if isinstance(ref, tuple) else [ref]
        scale = o.condition_scale(q)
        scale = list(scale) if isinstance(scale, tuple) else [scale]
        assert np.isnan(np.concatenate([np.asarray(r).reshape(len(q), -1) for r in ref],
                                       axis=1)[:-64]).any(), tag + ": no poisoned rows queried"
        paths = [("bucketed", N.ARB_Q_SORTED), ("direct", N.ARB_Q_DIRECT)]
        if prec == "mixed":
            paths.append(("table", N.ARB_Q_SORTED | N.ARB_Q_TABLE))
        else:
            paths.append(("fused", N.ARB_Q_SORTED | N.ARB_Q_FUSED))
        for pname, fl in paths:
            c, nv, g, cells = _gpu_query(t, q, fl)
            _compare(_outputs_for(mode, c, nv, g), ref, scale, cells, o.queryInds, prec,
                     tag + " " + pname)
        # single points: the first rows (near the bad node)
        for r in range(6):
            got = t.Query(q[r].copy())
            exp = o.Query(q[r].copy())
            got = got if isinstance(got, tuple) else (got,)
            exp = exp if isinstance(exp, tuple) else (exp,)
            for a, b in zip(got, exp):
                assert np.array_equal(np.isnan(np.atleast_1d(a)), np.isnan(np.atleast_1d(b))), \
                    tag + " point %d" % r
        t.close()


def test_nonfinite_table_columns(G):
    """Every coefficient of a poisoned cell is NaN, as in the reference for a
    NaN node; the neighbouring cells keep their finite coefficients."""
    n = (12, 11, 10)
    axes, vals = _field(3, n, 1, seed=5)
    node = (5, 4, 6)
    vb = [vals[0].copy()]
    vb[0][node] = np.nan
    rows = W.grid_rows(axes, vb)
    o = O.tricubic(rows.copy(), "quiet")
    o.allCoeffs()
    npos = [k - 3 for k in n]
    for prec in ("fp64", "mixed"):
        t = G.tricubic(rows.copy(), "quiet", precision=prec)
        tab = t.coefficient_table()[:-1, 0, :]
        cells = np.array([(x, y, z) for z in range(npos[2]) for y in range(npos[1])
                          for x in range(npos[0])])
        want = np.all((np.array(node)[None, :] >= cells) & (np.array(node)[None, :] <= cells + 3),
                      axis=1)
        ref = o.tables["n"][:, :-1].T
        assert np.isnan(ref[want]).all() and np.isfinite(ref[~want]).all()
        assert np.isnan(tab[want]).all(), prec
        assert np.isfinite(tab[~want]).all(), prec
        t.close()


def _field4(n, ncomp, seed):
    return _field(4, n, ncomp, seed)


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
@pytest.mark.parametrize("mode", ["scalar", "both"])
def test_nonfinite_nodes_4d(G, mode, prec):
    n = (11, 10, 9, 8)
    ncomp = 1 if mode == "scalar" else 3
    axes, vals = _field4(n, ncomp, seed=41 + len(mode))
    kw = {} if mode == "scalar" else {"mode": mode}
    rng = np.random.default_rng(19)
    for i, node in enumerate([(5, 4, 4, 3), (0, 9, 0, 7), (3, 3, 6, 4)]):
        bad = BAD[i % 3]
        comp = 0 if ncomp == 1 else (i + 1) % 3
        vb = [v.copy() for v in vals]
        vb[comp][node] = bad
        rows = W.grid_rows(axes, vb, np.random.default_rng(i))
        tag = "4d %s %s node %s comp%d=%s" % (prec, mode, node, comp, bad)
        print(tag, flush=True)
        t = G.quadcubic(rows.copy(), "quiet", precision=prec, **kw)
        o = O.quadcubic(rows.copy(), "quiet", **kw)
        q = _queries(axes, node, 3000, rng)
        ref = o.Query(q.copy())
        ref = list(ref) if isinstance(ref, tuple) else [ref]
        scale = o.condition_scale(q)
        scale = list(scale) if isinstance(scale, tuple) else [scale]
        for pname, fl in (("bucketed", N.ARB_Q_SORTED), ("direct", N.ARB_Q_DIRECT)):
            c, nv, g, cells = _gpu_query(t, q, fl)
            _compare(_outputs_for(mode, c, nv, g), ref, scale, cells, o.queryInds, prec,
                     tag + " " + pname)
        t.close()

"""Node-tile evaluation (mixed precision, 3D, bucketed path; modes 'both',
'vector', 'norm' and scalar fields).

Each query is evaluated from its cell's 4x4x4 node values with the Hermite
weights w_o(u) = sum_e M[e][o] u^e; a bucket's nodes are staged as FP32
differences from an FP64 polynomial model of each interpolant (linear for
the components, quadratic for |B|), which is added back per query in FP64.
Pinned here by:

  * the oracle (FP64 restatement of the reference), normwise 1e-6 as for
    every mixed result, on a dipole field;
  * exact reproduction: a field whose |B| is a quadratic polynomial is
    reproduced to FP64 rounding (value and gradient), because the weights
    reproduce quadratics and the model carries them in FP64; a linear
    component likewise;
  * invalidation: new node values (arb_field_set_values) rebuild the tiles
    (doubling every node doubles every output, bit for bit);
  * launch shapes: thread counts do not change the bits; other block shapes
    agree with the oracle.
"""
import ctypes

import numpy as np
import pytest

from arbinterp_amd import workloads as W
from oracle import arbinterp_oracle as O

pytestmark = pytest.mark.gpu
TOLMIX = 1e-6


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a ROCm GPU"
    from arbinterp_amd import ARBInterp as G
    return G


def _nodes_query(t, q, out_f64=True, **kw):
    """Bucketed default path (node tiles for mixed 3D fields)."""
    import torch
    from arbinterp_amd import _native as NN
    qd = torch.from_numpy(np.ascontiguousarray(q)).cuda()
    out = t.query_device(qd, out_f64=out_f64, cells=True, flags=NN.ARB_Q_SORTED, **kw)
    torch.cuda.synchronize()
    return [x.double().cpu().numpy() if x is not None else None for x in out]


@pytest.mark.parametrize("out_f64", [True, False])
def test_node_tiles_match_oracle(dev, out_f64):
    rows, q = W.dipole3d(seed=2, n=(48, 40, 36), nq=150_000)
    t = dev.tricubic(rows.copy(), "quiet", mode="both", precision="mixed")
    c, nv, g, cells = _nodes_query(t, q, out_f64=out_f64)
    info = t.field_info()
    assert info.tiles_bytes > 0 and all(info.block[a] > 0 for a in range(3))
    o = O.tricubic(rows.copy(), "quiet", mode="both")
    idx = np.random.default_rng(3).choice(len(q), 30_000, replace=False)
    co, no, go = o.Query(q[idx].copy())
    assert np.array_equal(cells[idx], o.queryInds)
    sc, sn, sg = o.condition_scale(q[idx])
    for got, ref, s, name in ((c[idx], co, sc, "comps"), (nv[idx, None], no, sn, "norm"),
                              (g[idx], go, sg, "grad")):
        mx, rms = O.rel_err_normwise(got, ref, s)
        print("node tiles %s (%s outputs): max %.3e rms %.3e"
              % (name, "FP64" if out_f64 else "FP32", mx, rms))
        assert mx <= TOLMIX, (name, mx, rms)
    t.close()


@pytest.mark.parametrize("mode", ["vector", "norm", "scalar"])
def test_node_tiles_other_modes(dev, mode):
    """'vector' (three components, 16-byte records, no gradient), 'norm' and
    scalar fields (one interpolant, 4-byte records, value and gradient).  The
    scalar is a dipole component, which crosses zero: there the FP32 table
    path misses the per-value bound (2.3e-5 measured on this field), while
    the node tiles' cubic model keeps the FP32 part small."""
    rows, q = W.dipole3d(seed=3, n=(44, 38, 34), nq=120_000)
    if mode == "scalar":
        rows = rows[:, :4].copy()
    kw = {} if mode == "scalar" else {"mode": mode}
    t = dev.tricubic(rows.copy(), "quiet", precision="mixed", **kw)
    out = _nodes_query(t, q)
    assert t.field_info().tiles_bytes > 0
    o = O.tricubic(rows.copy(), "quiet", **kw)
    idx = np.random.default_rng(9).choice(len(q), 20_000, replace=False)
    ref = o.Query(q[idx].copy())
    ref = ref if isinstance(ref, tuple) else (ref,)
    scales = o.condition_scale(q[idx])
    scales = scales if isinstance(scales, tuple) else (scales,)
    if mode == "vector":
        got = [out[0]]
    else:
        got = [out[1], out[2]]
    assert np.array_equal(out[3][idx], o.queryInds)
    for gk, r, s in zip(got, ref, scales):
        g2 = gk[idx].reshape(len(idx), -1)
        mx, rms = O.rel_err_normwise(g2, r.reshape(len(idx), -1), s.reshape(len(idx), -1))
        print("node tiles %s: max %.3e rms %.3e" % (mode, mx, rms))
        assert mx <= TOLMIX, (mode, mx)
    t.close()


def _quadratic_field(n=(26, 22, 19), h=(0.11, 0.09, 0.13), origin=(0.2, -0.4, 0.3)):
    axes = W.grid_axes(n, h, origin)
    X, Y, Z = np.meshgrid(*axes, indexing="ij")

    def quad(x, y, z):
        return (3.0 + 0.4 * x - 0.3 * y + 0.25 * z + 0.15 * x * x + 0.1 * y * y + 0.2 * z * z
                + 0.12 * x * y - 0.08 * x * z + 0.05 * y * z)

    def grad(x, y, z):
        return np.stack([0.4 + 0.3 * x + 0.12 * y - 0.08 * z,
                         -0.3 + 0.2 * y + 0.12 * x + 0.05 * z,
                         0.25 + 0.4 * z - 0.08 * x + 0.05 * y], axis=1)

    def lin(x, y, z):
        return 1.5 - 0.7 * x + 0.9 * y + 0.35 * z

    rows = W.grid_rows(axes, [quad(X, Y, Z), np.zeros_like(X), lin(X, Y, Z)],
                       np.random.default_rng(8))
    return rows, axes, quad, grad, lin


def test_node_tiles_reproduce_quadratics(dev):
    """Components: B = (q, 0, l) with q quadratic and l linear.  The linear
    component is reproduced exactly up to the node tiles' local coordinate
    (the components' model is linear and carried in FP64; the node tiles
    carry the cell-local coordinate as FP32 plus a correction in units of
    2^-32, a shift of at most 2^-32 cell per axis), the zero component is
    exactly zero, and q keeps its quadratic part in FP32 differences (a few
    2^-24 of that part).  |B|: with l = 0, |B| = q (q > 0 on the box), and
    value and gradient are reproduced to FP64 rounding plus that shift
    (|B|'s model is quadratic)."""
    rows, axes, quad, grad, lin = _quadratic_field()
    t = dev.tricubic(rows.copy(), "quiet", mode="both", precision="mixed")
    lo, hi = t._lo, t._hi
    h = np.array([t.hx, t.hy, t.hz])
    rng = np.random.default_rng(4)
    q = lo + (hi - lo) * rng.uniform(0.0, 1.0 - 1e-9, size=(120_000, 3))
    c, nv, g, cells = _nodes_query(t, q)
    x, y, z = q[:, 0], q[:, 1], q[:, 2]
    shift = 2.0 ** -31 * h                       # bound on the move of the point per axis
    ref_l = lin(x, y, z)
    slope_l = np.array([0.7, 0.9, 0.35])
    assert np.max(np.abs(c[:, 2] - ref_l)) <= (1e-13 * np.max(np.abs(ref_l))
                                                + 1.01 * np.dot(slope_l, shift))
    assert np.all(c[:, 1] == 0.0)
    ref_q = quad(x, y, z)
    assert np.max(np.abs(c[:, 0] - ref_q)) <= 1e-6 * np.max(np.abs(ref_q))
    t.close()

    rows2 = rows.copy()
    rows2[:, 5] = 0.0
    t = dev.tricubic(rows2, "quiet", mode="both", precision="mixed")
    c, nv, g, cells = _nodes_query(t, q)
    ref_g = grad(x, y, z)
    gmax = np.abs(ref_g).max(axis=0)             # per-axis bound of |dq/dx_a|
    hess = np.array([[0.3, 0.12, -0.08], [0.12, 0.2, 0.05], [-0.08, 0.05, 0.4]])
    assert np.max(np.abs(nv - ref_q)) <= (1e-13 * np.max(np.abs(ref_q))
                                          + 1.01 * np.dot(gmax, shift))
    gerr = np.abs(g - ref_g).max(axis=0)
    assert np.all(gerr <= 1e-12 * np.max(np.abs(ref_g)) + 1.01 * np.abs(hess) @ shift), gerr
    t.close()


def test_node_tiles_follow_new_values(dev):
    import torch
    from arbinterp_amd import _native as NN
    rows, q = W.dipole3d(seed=5, n=(30, 27, 25), nq=100_000)
    t = dev.tricubic(rows.copy(), "quiet", mode="both", precision="mixed")
    a = _nodes_query(t, q)
    info = t.field_info()
    vals = torch.empty((info.ninterp, info.nnodes), dtype=torch.float64, device="cuda")
    lib = NN.lib()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    NN.check(lib.arb_field_values(t._h, ctypes.c_void_p(vals.data_ptr()), stream), "values")
    vals *= 2.0
    NN.check(lib.arb_field_set_values(t._h, ctypes.c_void_p(vals.data_ptr()), stream), "set")
    b = _nodes_query(t, q)
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(np.nan_to_num(2.0 * x, nan=5.0), np.nan_to_num(y, nan=5.0))
    assert np.array_equal(a[3], b[3])
    t.close()


def test_node_tiles_launch_shapes(dev):
    rows, q = W.dipole3d(seed=2, n=(33, 29, 27), nq=120_000)
    t = dev.tricubic(rows.copy(), "quiet", mode="both", precision="mixed")
    o = O.tricubic(rows.copy(), "quiet", mode="both")
    idx = np.random.default_rng(6).choice(len(q), 15_000, replace=False)
    ref = o.Query(q[idx].copy())
    scales = o.condition_scale(q[idx])
    for block in ((4, 4, 2), (2, 2, 2), (4, 4, 4), (2, 2, 1)):
        outs = []
        for threads in (64, 256):
            t.set_tuning(block, threads)
            outs.append(_nodes_query(t, q))
            assert tuple(t.field_info().block[a] for a in range(3)) == block
        for x, y in zip(*outs):
            assert np.array_equal(np.nan_to_num(x, nan=5.0), np.nan_to_num(y, nan=5.0)), block
        c, nv, g, _ = outs[0]
        for got, r, s in ((c[idx], ref[0], scales[0]), (nv[idx, None], ref[1], scales[1]),
                          (g[idx], ref[2], scales[2])):
            mx, _ = O.rel_err_normwise(got, r, s)
            assert mx <= TOLMIX, (block, mx)
    t.close()


def test_node_tiles_side_outputs(dev):
    """Rows outside the box, NaN rows and IntMax rows on the node-tile paths:
    the caller's rows are NaN-filled as the reference does, cell indices are
    the oracle's, the direct and bucketed node-tile paths give the same bits,
    and a single point past the sentinel raises IndexError."""
    import torch
    from arbinterp_amd import _native as NN
    rows, q = W.dipole3d(seed=7, n=(26, 23, 21), nq=20_000)
    t = dev.tricubic(rows.copy(), "quiet", mode="both", precision="mixed")
    lo, hi = np.array(t._lo), np.array(t._hi)
    rng = np.random.default_rng(12)
    extra = rng.uniform(lo - 0.3 * (hi - lo), hi + 0.3 * (hi - lo), size=(2000, 3))
    edge = np.tile((lo + hi) / 2, (4, 1))
    edge[0, 0] = hi[0]
    edge[1, 1] = hi[1]
    edge[2, 0] = lo[0]
    q = np.concatenate([q, extra, edge, np.full((2, 3), np.nan)])
    q = q[rng.permutation(len(q))]
    o = O.tricubic(rows.copy(), "quiet", mode="both")
    qa = q.copy()
    ref = o.Query(qa)                       # NaN-fills rows outside the box
    outs = {}
    for name, fl in (("direct", NN.ARB_Q_DIRECT), ("sorted", NN.ARB_Q_SORTED),
                     ("cell", NN.ARB_Q_SORTED | NN.ARB_Q_SORT_CELL)):
        qd = torch.from_numpy(np.concatenate([q, np.ones((len(q), 1))], axis=1)).cuda()
        c, nv, g, cells = t.query_device(qd, nan_rows=True, cells=True, out_f64=True, flags=fl)
        torch.cuda.synchronize()
        qm = qd.cpu().numpy()
        assert np.array_equal(np.isnan(qm[:, :3]), np.isnan(qa)), name
        oob = np.isnan(qa).any(axis=1) & ~np.isnan(q).any(axis=1)
        assert oob.sum() > 100 and np.all(np.isnan(qm[oob, 3])), name   # whole rows
        assert np.array_equal(cells.cpu().numpy(), o.queryInds), name
        outs[name] = [x.cpu().numpy() for x in (c, nv, g)]
    assert t.field_info().tiles_bytes > 0
    for name in ("sorted", "cell"):
        for a, b in zip(outs["direct"], outs[name]):
            assert np.array_equal(np.nan_to_num(a, nan=9.0), np.nan_to_num(b, nan=9.0)), name
    for a, r in zip(outs["direct"], ref):
        assert np.array_equal(np.isnan(a.reshape(len(q), -1)).any(axis=1),
                              np.isnan(r.reshape(len(q), -1)).any(axis=1))
    with pytest.raises(IndexError):           # the far corner, as a single point
        t.Query(hi.copy())
    t.close()


@pytest.mark.parametrize("mode", ["both", "scalar"])
def test_node_tiles_tiny_grids(dev, mode):
    """Grids of 4 and 5 nodes per axis (one and two cells): the tile model's
    middle node is clamped into the grid and the third differences along an
    axis of 4 nodes are left out; results still match the oracle."""
    rng = np.random.default_rng(31)
    for n in ((4, 5, 6), (5, 4, 4), (4, 4, 4)):
        axes = W.grid_axes(n, (0.3, 0.25, 0.2), (0.1, -0.2, 0.4))
        X, Y, Z = np.meshgrid(*axes, indexing="ij")
        vals = [np.sin(2.0 * X + 0.5) * np.cos(1.5 * Y) + 0.3 * Z,
                np.cos(X - Y) + 0.2, 0.4 * X * Y - np.sin(Z)]
        if mode == "scalar":
            vals = vals[:1]
        rows = W.grid_rows(axes, vals, rng)
        kw = {} if mode == "scalar" else {"mode": "both"}
        t = dev.tricubic(rows.copy(), "quiet", precision="mixed", **kw)
        o = O.tricubic(rows.copy(), "quiet", **kw)
        lo, hi = np.array(t._lo), np.array(t._hi)
        q = lo + (hi - lo) * rng.uniform(0.0, 1.0 - 1e-9, size=(3000, 3))
        out = _nodes_query(t, q)
        assert t.field_info().tiles_bytes > 0, n
        ref = o.Query(q.copy())
        ref = ref if isinstance(ref, tuple) else (ref,)
        got = [x for x in out[:3] if x is not None]
        scales = o.condition_scale(q)
        scales = scales if isinstance(scales, tuple) else (scales,)
        for gk, r, s in zip(got, ref, scales):
            m = len(q)
            mx, _ = O.rel_err_normwise(gk.reshape(m, -1), r.reshape(m, -1), s.reshape(m, -1),
                                       floor=1.0)
            assert mx <= 2e-7, (n, mode, mx)
        t.close()


def test_node_tiles_need_no_table(dev):
    """Node-tile queries leave the coefficient table unallocated: the
    device memory the field takes through creation and a batch of queries
    is well below the table's size; asking for the table (allCoeffs) then
    allocates it."""
    import torch
    rows, q = W.dipole3d(seed=2, n=(96, 96, 96), nq=300_000)
    qd = torch.from_numpy(q).cuda()
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    t = dev.tricubic(rows.copy(), "quiet", mode="both", precision="mixed")
    out = t.query_device(qd)
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    info = t.field_info()
    assert info.tiles_bytes > 0
    table = int(info.table_bytes)                   # 93^3 cells x 1 KB: about 0.8 GB
    assert free0 - free1 < table // 2, (free0 - free1, table)
    t.allCoeffs()
    torch.cuda.synchronize()
    free2, _ = torch.cuda.mem_get_info()
    assert free1 - free2 >= table * 9 // 10, (free1 - free2, table)
    del out
    t.close()

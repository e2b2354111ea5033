"""GPU parity: the drop-in (libarbinterp.so kernels) against the reference's
golden vectors and the CPU oracle.  Bar for the FP64 plain path: identical cell
indices, NaN patterns, mutated query arrays and exceptions; values and
gradients within 1e-12 under oracle.rel_err (|d| / max(|ref|, 1e-4 S))."""
import numpy as np
import pytest

from arbinterp_amd import workloads as W
from oracle import arbinterp_oracle as O
from tests.parity import compare_outputs, run_query

pytestmark = pytest.mark.gpu
TOL64 = 1e-12
TOLMIX = 1e-6      # normwise per query (oracle.rel_err_normwise)


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a ROCm GPU"
    from arbinterp_amd import ARBInterp as G
    return G


def test_scalar3d_gaussians_golden(dev, golden):
    g = golden("scalar3d_gaussians")
    t = dev.tricubic(g["field"].copy(), "quiet")
    o = O.tricubic(g["field"].copy(), "quiet")
    res, qa = run_query(t, g["query"])
    compare_outputs(res, qa, g, "", o, g["query"], tol=TOL64)
    assert np.array_equal(t.queryInds, g["queryInds"])


@pytest.mark.parametrize("mode", ["vector", "norm", "both", "bogus", "default"])
def test_dipole3d_modes_golden(dev, golden, mode):
    g = golden("dipole3d_modes")
    kw = {} if mode == "default" else {"mode": mode}
    t = dev.tricubic(g["field"].copy(), "quiet", **kw)
    o = O.tricubic(g["field"].copy(), "quiet", **kw)
    res, qa = run_query(t, g["query"])
    compare_outputs(res, qa, g, mode + "_", o, g["query"], tol=TOL64)
    assert np.array_equal(t.queryInds, g[mode + "_queryInds"])


def test_moving4d_scalar_golden(dev, golden):
    g = golden("moving4d_scalar")
    t = dev.quadcubic(g["field"].copy(), "quiet")
    o = O.quadcubic(g["field"].copy(), "quiet")
    res, qa = run_query(t, g["query"])
    compare_outputs(res, qa, g, "", o, g["query"], tol=TOL64)
    assert np.array_equal(t.queryInds, g["queryInds"])


@pytest.mark.parametrize("mode", ["vector", "norm", "both", "default"])
def test_vector4d_modes_golden(dev, golden, mode):
    g = golden("vector4d_modes")
    kw = {} if mode == "default" else {"mode": mode}
    t = dev.quadcubic(g["field"].copy(), "quiet", **kw)
    o = O.quadcubic(g["field"].copy(), "quiet", **kw)
    res, qa = run_query(t, g["query"])
    compare_outputs(res, qa, g, mode + "_", o, g["query"], tol=TOL64)
    assert np.array_equal(t.queryInds, g[mode + "_queryInds"])


def test_coefficient_tables_golden(dev, golden):
    g = golden("coeff_tables")
    t = dev.tricubic(g["field3"].copy(), "quiet", mode="both")
    tab = t.coefficient_table()                 # (nc+1, 4, 64)
    o = O.tricubic(g["field3"].copy(), "quiet", mode="both")
    assert np.array_equal(t.basePointInds, g["t3_basePointInds"])
    c = g["t3_consts"]
    assert [t.hx, t.hy, t.hz, t.xIntMin, t.yIntMin, t.zIntMin, t.xIntMax, t.yIntMax,
            t.zIntMax, t.nc] == list(c)
    off = O.neighbour_offsets(o.p.n, 3)
    for k, name in enumerate(("alphax", "alphay", "alphaz", "alphan")):
        ref = g["t3_" + name]                    # (64, nc+1)
        got = tab[:, k, :].T
        assert np.array_equal(np.isnan(got), np.isnan(ref))
        vals = o.interps["xyzn"[k]]
        scale = np.abs(O.make_A(3)) @ np.abs(vals[o.p.base[None, :] + off[:, None]])
        assert (np.abs(got[:, :-1] - ref[:, :-1]) <= 1e-15 * scale).all(), name
    t4 = dev.quadcubic(g["field4"].copy(), "quiet")
    o4 = O.quadcubic(g["field4"].copy(), "quiet")
    tab4 = t4.coefficient_table()
    ref = g["t4_alphan"]
    got = tab4[:, 0, :].T
    off4 = O.neighbour_offsets(o4.p.n, 4)
    scale = np.abs(O.make_A(4)) @ np.abs(o4.interps["n"][o4.p.base[None, :] + off4[:, None]])
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    assert (np.abs(got[:, :-1] - ref[:, :-1]) <= 1e-15 * scale).all()


def test_mixed_coefficient_table_bound(dev):
    """Mixed-precision 3D table against the oracle's FP64 alpha: the table holds
    alpha of the node values quantised to 2^(E-32) (E: exponent of the cell's
    largest |F - f0|) rounded once to FP32, so
        |got - ref| <= 2^-24 |ref| + 2^(E-32) sum_k |A_ik|."""
    rows, _ = W.dipole3d(seed=2, n=(21, 18, 15), nq=10)
    t = dev.tricubic(rows.copy(), "quiet", mode="both", precision="mixed")
    tab = t.coefficient_table().astype(np.float64)     # (nc+1, 4, 64)
    o = O.tricubic(rows.copy(), "quiet", mode="both")
    A = O.make_A(3)
    off = O.neighbour_offsets(o.p.n, 3)
    rowsum = np.abs(A).sum(axis=1)
    for k, name in enumerate("xyzn"):
        F = o.interps[name][o.p.base[None, :] + off[:, None]]     # (64, nc)
        ref = A @ F
        r = np.abs(F - F[21][None, :]).max(axis=0)
        E = np.frexp(r)[1]
        bound = 2.0 ** -24 * np.abs(ref) + np.ldexp(rowsum[:, None], E[None, :] - 32)
        got = tab[:-1, k, :].T
        assert (np.abs(got - ref) <= bound).all(), name


def test_edge_battery_golden(dev, golden):
    g = golden("edge_battery")
    rows, _ = W.scalar3d_gaussians(seed=1)
    vrows, _ = W.dipole3d(seed=2, n=(14, 12, 10), nq=10)
    for i in range(int(g["ncases"])):
        pre = "c%03d_" % i
        fld = rows if bool(g[pre + "scalar"]) else vrows
        mode = str(g[pre + "mode"])
        kw = {} if mode == "none" else {"mode": mode}
        o = O.tricubic(fld.copy(), "quiet", **kw)
        t = dev.tricubic(fld.copy(), "quiet", **kw)
        res, qa = run_query(t, g[pre + "batch"])
        compare_outputs(res, qa, g, pre + "b_", o, g[pre + "batch"], tol=TOL64, label=pre + mode)
        p = g[pre + "batch"][0]
        t2 = dev.tricubic(fld.copy(), "quiet", **kw)
        res1, qa1 = run_query(t2, p)
        compare_outputs(res1, qa1, g, pre + "s_", o, p[None, :], tol=TOL64, label=pre + mode)


# A float32 field queried with float32 positions: the reference reduces
# cells in float32 and raises the float32 local coordinates to their powers
# in float32 (numpy promotion); so does ARB_Q_REDUCE_F32, with the float32
# cube rounded to nearest.  numpy's own vectorised float32 power differs from
# the nearest float32 in about one case in five (an ulp of u^3), so values
# agree to 1e-8 of the condition-floored scale, not to FP64 rounding
# (measured 6e-10 value, 2e-8 gradient; 5.5e-8 and 2.8e-6 with an FP64
# reduction).  Cell indices and NaN rows are identical.
TOL_F32F32 = 1e-7
# mixed precision near the zeros of a fast-varying scalar field (floored
# normwise metric): node tiles and the FP32 coefficient table
TOL_NODES_ZC = 5e-6
TOL_TABLE_ZC = 5e-5


@pytest.mark.parametrize("tag", ["f64f32", "f32f64", "f32f32"])
def test_dtype_promotion_golden(dev, golden, tag):
    g = golden("dtype_promotion")
    t = dev.tricubic(g[tag + "_field"].copy(), "quiet")
    o = O.tricubic(g[tag + "_field"].copy(), "quiet")
    res, qa = run_query(t, g[tag + "_query"])
    compare_outputs(res, qa, g, tag + "_", o, g[tag + "_query"],
                    tol=TOL_F32F32 if tag == "f32f32" else TOL64)
    assert np.array_equal(t.queryInds, g[tag + "_queryInds"])
    if tag == "f32f32":
        # single float32 points take the float32 reduction too (an
        # out-of-box point makes Query2 raise TypeError in both)
        q32 = g[tag + "_query"]
        for r in range(0, len(q32), 37):
            try:
                pv = t.Query(q32[r].copy())
            except TypeError:
                pv = None
            try:
                ov = o.Query(q32[r].copy())
            except TypeError:
                ov = None
            assert (pv is None) == (ov is None), r
            if pv is not None:
                assert t.queryInd == o.queryInd == g[tag + "_queryInds"][r], r
                assert np.allclose(pv[0], ov[0], rtol=1e-6, atol=1e-9), r
                assert np.allclose(pv[1], ov[1], rtol=1e-5, atol=1e-7), r


def test_reduce_f32_at_cell_edges(dev):
    """Points one float32 ulp below a cell edge: the float32 reduction may
    round iu up to the edge and give the next cell, as numpy does; the cell
    index must follow the float32 arithmetic exactly."""
    import torch
    from arbinterp_amd import _native as NN
    rows, _ = W.scalar3d_gaussians(seed=7, n=(12, 11, 10), nq=10)
    f32 = rows.astype(np.float32)
    t = dev.tricubic(f32.copy(), "quiet")
    o = O.tricubic(f32.copy(), "quiet")
    rng = np.random.default_rng(2)
    cells = rng.integers(1, 6, size=(400, 3))
    edge = np.stack([o.p.int_min[a] + cells[:, a] * o.p.h[a] for a in range(3)], axis=1)
    q = np.nextafter(edge.astype(np.float32), np.float32(-np.inf))
    q[::2] = edge[::2].astype(np.float32)
    qd = torch.from_numpy(q.astype(np.float64)).cuda()
    _, _, _, cl = t.query_device(qd, cells=True, flags=NN.ARB_Q_REDUCE_F32)
    torch.cuda.synchronize()
    o.Query(q.copy())
    assert np.array_equal(cl.cpu().numpy(), o.queryInds)


def test_mixed_precision_error_bound(dev):
    rows, q = W.dipole3d(seed=2, n=(40, 36, 32), nq=50_000)
    t = dev.tricubic(rows.copy(), "quiet", mode="both", precision="mixed")
    o = O.tricubic(rows.copy(), "quiet", mode="both")
    c, nv, gr = t.Query(q.copy())
    co, no, go = o.Query(q.copy())
    sc, sn, sg = o.condition_scale(q)
    for got, ref, s, name in ((c, co, sc, "comps"), (nv, no, sn, "norm"), (gr, go, sg, "grad")):
        mx, rms = O.rel_err_normwise(got, ref, s)
        print("mixed %s: max %.3e rms %.3e" % (name, mx, rms))
        assert mx <= TOLMIX, (name, mx, rms)


def test_device_path_matches_host_path(dev):
    import torch
    rows, q = W.scalar3d_gaussians(seed=1)
    t = dev.tricubic(rows.copy(), "quiet")
    vh, gh = t.Query(q.copy())
    qd = torch.from_numpy(np.concatenate([q, np.ones((len(q), 2))], axis=1)).cuda()
    _, nd, gd = t.query_device(qd, nan_rows=True)
    torch.cuda.synchronize()
    assert np.array_equal(np.nan_to_num(nd.cpu().numpy(), nan=3.0), np.nan_to_num(vh[:, 0], nan=3.0))
    assert np.array_equal(np.nan_to_num(gd.cpu().numpy(), nan=3.0), np.nan_to_num(gh, nan=3.0))
    qa = q.copy()
    o = O.tricubic(rows.copy(), "quiet")
    o.Query(qa)
    oob = np.isnan(qa).any(axis=1) & ~np.isnan(q).any(axis=1)
    qd_h = qd.cpu().numpy()
    assert np.isnan(qd_h[oob]).all() and not np.isnan(qd_h[~oob]).any()


def test_deterministic_repeat(dev):
    rows, q = W.dipole3d(seed=2, n=(30, 28, 26), nq=20_000)
    t = dev.tricubic(rows.copy(), "quiet", mode="both", precision="mixed")
    a = t.Query(q.copy())
    b = t.Query(q.copy())
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_full_size_dipole_sample(dev):
    """128^3 vector field, mode 'both': 10^6 queries on the GPU; 20k of them
    checked against the oracle; every row in range and finite.  At this size
    the FP64 buckets are 4x4x2 cells (66.5 KB of LDS per workgroup), and the
    fused kernel (another 14 KB) must give the same bits."""
    import torch
    from arbinterp_amd import _native as NN
    rows, q = W.dipole3d(seed=2, nq=1_000_000)
    t = dev.tricubic(rows, "quiet", mode="both")
    qd = torch.from_numpy(q).cuda()
    c, nv, gr, cells = t.query_device(qd, cells=True)
    fused = t.query_device(qd, flags=NN.ARB_Q_SORTED | NN.ARB_Q_FUSED)
    torch.cuda.synchronize()
    assert int(cells.max()) < t.nc and torch.isfinite(nv).all()
    for a, b in zip((c, nv, gr), fused[:3]):
        assert torch.equal(a, b)
    o = O.tricubic(rows, "quiet", mode="both")
    idx = np.random.default_rng(0).choice(len(q), 20_000, replace=False)
    co, no, go = o.Query(q[idx].copy())
    assert np.array_equal(cells.cpu().numpy()[idx], o.queryInds)
    sc, sn, sg = o.condition_scale(q[idx])
    for got, ref, s in ((c.cpu().numpy()[idx], co, sc), (nv.cpu().numpy()[idx, None], no, sn),
                        (gr.cpu().numpy()[idx], go, sg)):
        mx, _ = O.rel_err(got, ref, s)
        assert mx <= TOL64, mx


def _nan_equal(a, b):
    a = a.cpu().numpy() if hasattr(a, "cpu") else a
    b = b.cpu().numpy() if hasattr(b, "cpu") else b
    return np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(a[~np.isnan(a)],
                                                                       b[~np.isnan(b)])


@pytest.mark.parametrize("case", ["3d_both_mixed", "3d_both_fp64", "3d_scalar", "4d_scalar",
                                  "4d_both_mixed", "3d_vector_tuned", "3d_scalar_mixed",
                                  "3d_vector_mixed_tuned"])
def test_query_paths_bitwise_identical(dev, case):
    """Direct, bucketed-LDS and bucketed-global paths give identical bits,
    including rows outside the box, NaN rows, IntMax rows, partial blocks.
    For mixed 3D the fused path forms coefficients in the query kernel and
    the others read the table: the same bits are required.  Mixed 3D fields
    evaluate direct and bucketed queries from node tiles by default: those
    paths give the same bits as each other and agree with the table family to
    FP32 rounding (checked against the oracle)."""
    import torch
    from arbinterp_amd import _native as NN
    rng = np.random.default_rng(5)
    if case.startswith("3d"):
        rows, q = W.dipole3d(seed=2, n=(23, 19, 14), nq=30_000)
        if case.startswith("3d_scalar"):
            rows = rows[:, :4].copy()
    else:
        rows, q = W.moving4d(seed=3, n=(11, 9, 8, 7), nq=30_000)
        if case == "4d_both_mixed":
            extra = np.stack([np.sin(rows[:, 0] * 3 + rows[:, 3]), np.cos(rows[:, 1] * 2),
                              rows[:, 4]], axis=1)
            rows = np.concatenate([rows[:, :4], extra], axis=1)
    cls = dev.tricubic if case.startswith("3d") else dev.quadcubic
    kw = {}
    if "both" in case:
        kw["mode"] = "both"
    if "vector" in case:
        kw["mode"] = "vector"
    prec = "mixed" if "mixed" in case else "fp64"
    t = cls(rows, "quiet", precision=prec, **kw)
    if case in ("3d_vector_tuned", "3d_vector_mixed_tuned"):
        t.set_tuning((2, 2, 8), 128)
    d = t.NDIM
    lo = np.array([getattr(t, "xyzt"[a] + "IntMin") for a in range(d)])
    hi = np.array([getattr(t, "xyzt"[a] + "IntMax") for a in range(d)])
    extra = rng.uniform(lo - 0.2 * (hi - lo), hi + 0.2 * (hi - lo), size=(3000, d))
    edge = np.tile((lo + hi) / 2, (6, 1))
    for a in range(min(d, 3)):
        edge[2 * a, a] = hi[a]
        edge[2 * a + 1, a] = lo[a]
    q = np.concatenate([q, extra, edge, np.full((3, d), np.nan)], axis=0)
    q = q[rng.permutation(len(q))]
    qd = torch.from_numpy(np.ascontiguousarray(q)).cuda()
    outs = {}
    for name, fl in (("direct", NN.ARB_Q_DIRECT), ("lds", NN.ARB_Q_SORTED),
                     ("cell", NN.ARB_Q_SORTED | NN.ARB_Q_SORT_CELL),
                     ("table", NN.ARB_Q_SORTED | NN.ARB_Q_TABLE),
                     ("fused", NN.ARB_Q_SORTED | NN.ARB_Q_FUSED),
                     ("atomic", NN.ARB_Q_SORTED | NN.ARB_Q_GROUP_ATOMIC),
                     ("global", NN.ARB_Q_SORTED_GLOBAL)):
        # fresh buffers holding a marker, so a row that a path skips (or a
        # sort that loses a row) cannot pass on stale memory
        buf = t.query_device(qd.clone(), cells=True, flags=NN.ARB_Q_DIRECT)
        torch.cuda.synchronize()
        for x in buf:
            if x is not None:
                x.fill_(-7 if x.dtype == torch.int64 else 3.0e30)
        outs[name] = t.query_device(qd.clone(), cells=True, flags=fl, out=buf)
    torch.cuda.synchronize()
    # mixed 3D fields evaluate direct and bucketed queries from node tiles
    # where the block shape has a node-tile kernel (not for the (2, 2, 8)
    # tuning); the table-reading paths form the other family
    node_tiles = t.field_info().tiles_bytes > 0
    assert node_tiles == (case in ("3d_both_mixed", "3d_scalar_mixed"))
    ref_of = {"lds": "direct", "cell": "direct", "table": "direct", "fused": "direct",
              "atomic": "direct", "global": "direct"}
    if node_tiles:
        ref_of.update({"fused": "table", "atomic": "table", "global": "table"})
        del ref_of["table"]
    for name, rn in ref_of.items():
        for a, b in zip(outs[rn], outs[name]):
            if a is None:
                assert b is None
                continue
            assert _nan_equal(a, b), (case, name)
    if node_tiles:
        assert torch.equal(outs["table"][3], outs["direct"][3])
        # the same NaN rows as the table family, and the oracle's values on
        # every row inside the box: within 2e-7 of the condition scale (the
        # FP32 bound), and under the floored metric within the stated bound
        # of each path.  A scalar field that crosses zero and varies fast
        # across a cell (one dipole component on this coarse 23 x 19 x 14
        # grid) is the hard case for FP32 evaluation: there the node tiles
        # are bounded by 5e-6 (measured 3.3e-6) and the FP32 table by 5e-5
        # (2.5e-5); the 'both' field holds 1e-6 on both paths
        for a, b in zip(outs["table"][:3], outs["lds"][:3]):
            if a is not None:
                assert torch.equal(torch.isnan(a), torch.isnan(b)), case
        cells = outs["direct"][3].cpu().numpy()
        rows_in = np.where((cells < t.nc) & ~np.isnan(q).any(axis=1))[0]
        o = O.tricubic(rows, "quiet", **kw)
        ref = o.Query(q[rows_in].copy())
        ref = ref if isinstance(ref, tuple) else (ref,)
        scales = o.condition_scale(q[rows_in])
        scales = scales if isinstance(scales, tuple) else (scales,)
        got = [b for b in outs["lds"][:3] if b is not None]
        tab = [b for b in outs["table"][:3] if b is not None]
        assert len(got) == len(tab) == len(ref) == len(scales)
        m = len(rows_in)
        for k, (b, bt, r, s) in enumerate(zip(got, tab, ref, scales)):
            b = b.double().cpu().numpy().reshape(len(q), -1)[rows_in]
            bt = bt.double().cpu().numpy().reshape(len(q), -1)[rows_in]
            r, s = r.reshape(m, -1), s.reshape(m, -1)
            mx, _ = O.rel_err_normwise(b, r, s)
            mt, _ = O.rel_err_normwise(bt, r, s)
            ms, _ = O.rel_err_normwise(b, r, s, floor=1.0)
            print("%s node tiles k=%d: %.3e (table %.3e), relative to S %.3e" % (case, k, mx, mt, ms))
            assert ms <= 2e-7, (case, k, ms)
            zero_crossing = case == "3d_scalar_mixed"
            assert mx <= (TOL_NODES_ZC if zero_crossing else TOLMIX), (case, k, mx)
            assert mt <= (TOL_TABLE_ZC if zero_crossing else TOLMIX), (case, k, mt)


def test_coefficient_save_load_roundtrip(dev, tmp_path):
    rows, q = W.dipole3d(seed=2, n=(20, 18, 16), nq=5000)
    from arbinterp_amd import _native as NN
    for prec in ("fp64", "mixed"):
        t = dev.tricubic(rows.copy(), "quiet", mode="both", precision=prec)
        # a loaded table is authoritative (queries read it): compare with the
        # table path (mixed 3D queries otherwise use node tiles)
        t.default_flags = NN.ARB_Q_TABLE
        a = t.Query(q.copy())
        path = str(tmp_path / ("tab_%s.npz" % prec))
        t.save_coefficients(path)
        t2 = dev.tricubic(rows.copy(), "quiet", mode="both", precision=prec)
        t2.load_coefficients(path)
        b = t2.Query(q.copy())
        for x, y in zip(a, b):
            assert np.array_equal(np.nan_to_num(x, nan=7.0), np.nan_to_num(y, nan=7.0))
        other = dev.tricubic(rows.copy(), "quiet", mode="norm", precision=prec)
        with pytest.raises(ValueError):
            other.load_coefficients(path)


@pytest.mark.parametrize("case", ["3d_both_mixed", "3d_scalar_fp64", "3d_scalar_mixed", "3d_vector",
                                  "4d_both_fp64", "4d_scalar_mixed"])
def test_packed_records_match_arrays(dev, case):
    import torch
    from arbinterp_amd import _native as NN
    if case.startswith("3d"):
        rows, q = W.dipole3d(seed=2, n=(21, 19, 17), nq=70_000)
        if "scalar" in case:
            rows = rows[:, :4].copy()
        cls = dev.tricubic
    else:
        rows, q = W.moving4d(seed=3, n=(10, 9, 8, 7), nq=70_000)
        if "both" in case:
            extra = np.stack([np.sin(rows[:, 0]), np.cos(rows[:, 1]), rows[:, 4]], axis=1)
            rows = np.concatenate([rows[:, :4], extra], axis=1)
        cls = dev.quadcubic
    kw = {"mode": "both"} if "both" in case else ({"mode": "vector"} if "vector" in case else {})
    t = cls(rows, "quiet", precision="mixed" if "mixed" in case else "fp64", **kw)
    qd = torch.from_numpy(np.concatenate([q, q[:100] * 1.3])).cuda()
    for fl in (NN.ARB_Q_DIRECT, NN.ARB_Q_SORTED):
        ref = t.query_device(qd, flags=fl)
        rec = t.query_packed(qd, flags=fl)
        for a, b in zip(ref, t.unpack(rec)):
            if a is None:
                assert b is None
                continue
            assert _nan_equal(a, b.contiguous()), (case, fl)
    # the bucketed path splits its packed workspace records into the arrays:
    # the same bits as the direct path, also for a later, smaller call that
    # reuses the workspace, and with FP64 outputs of a mixed field
    for f64 in ((False, True) if "mixed" in case else (False,)):
        for m in (len(qd), 40_000):
            a = t.query_device(qd[:m], flags=NN.ARB_Q_DIRECT, out_f64=f64)
            b = t.query_device(qd[:m], flags=NN.ARB_Q_SORTED, out_f64=f64)
            for x, y in zip(a, b):
                if x is None:
                    assert y is None
                    continue
                assert _nan_equal(x, y), (case, f64, m)


def _point_cases(golden):
    g = golden("vector4d_modes")
    rows3, _ = W.dipole3d(seed=2, n=(20, 18, 16), nq=10)
    srows, _ = W.scalar3d_gaussians(seed=1)
    cases = [("tricubic", srows, None)]
    cases += [("tricubic", rows3, m) for m in ("both", "vector", "norm")]
    cases += [("quadcubic", g["field"], m) for m in ("both", "vector", "norm")]
    return cases


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_single_point_matches_batch(dev, golden, precision):
    """Query(point) runs the point mailbox entry; each call must give the batch
    path's row bit for bit (both evaluate the same cell with FP64 outputs),
    including points on the box faces, and the cell index."""
    for cls, fld, mode in _point_cases(golden):
        kw = {"precision": precision} if mode is None else {"mode": mode, "precision": precision}
        t = getattr(dev, cls)(fld.copy(), "quiet", **kw)
        d = t.NDIM
        lo, hi = np.array(t._lo[:d]), np.array(t._hi[:d])
        rng = np.random.default_rng(17)
        pts = lo + (hi - lo) * rng.uniform(size=(64, d))
        pts[0] = lo
        pts[1, 0] = hi[0]
        # the far corner reduces past the sentinel: IndexError, as the oracle
        with pytest.raises(IndexError):
            t.Query(hi.copy())
        batch = t.Query(pts.copy())
        batch = batch if isinstance(batch, tuple) else (batch,)
        inds = t.queryInds.copy()
        for i, p in enumerate(pts):
            one = t.Query(p.copy())
            one = one if isinstance(one, tuple) else (one,)
            assert len(one) == len(batch)
            for a, b in zip(one, batch):
                assert np.array_equal(np.asarray(a, dtype=np.float64).ravel(),
                                      np.asarray(b[i], dtype=np.float64).ravel()), \
                    (cls, mode, precision, i)
            assert t.queryInd == inds[i], (cls, mode, precision, i)
        t.close()


@pytest.mark.parametrize("dim", [3, 4])
def test_empty_query_batches(dev, golden, dim):
    """A (0, d) batch returns empty outputs of the oracle's shapes (numpy's
    own behaviour in the reference's rQuery), and the device path returns
    empty tensors without launching anything."""
    import torch
    if dim == 3:
        rows, _ = W.dipole3d(seed=2, n=(12, 11, 10), nq=10)
        cls, ocls = dev.tricubic, O.tricubic
    else:
        rows = golden("vector4d_modes")["field"]
        cls, ocls = dev.quadcubic, O.quadcubic
    fields = [(rows[:, :dim + 1].copy(), {})] + [(rows, {"mode": m}) for m in ("both", "vector", "norm")]
    for fld, kw in fields:
        for prec in ("fp64", "mixed"):
            t = cls(fld.copy(), "quiet", precision=prec, **kw)
            o = ocls(fld.copy(), "quiet", **kw)
            got = t.Query(np.zeros((0, dim)))
            ref = o.Query(np.zeros((0, dim)))
            got = got if isinstance(got, tuple) else (got,)
            ref = ref if isinstance(ref, tuple) else (ref,)
            assert [g.shape for g in got] == [r.shape for r in ref], (kw, prec)
            assert t.queryInds.shape == (0,)
            out = t.query_device(torch.zeros((0, dim), dtype=torch.float64, device="cuda"))
            assert all(x is None or x.shape[0] == 0 for x in out)
            t.close()

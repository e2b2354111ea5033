"""The oracle against golden vectors written by the reference itself
(tests/golden/make_golden.py, run where the reference is mounted)."""
import numpy as np
import pytest

from oracle import arbinterp_oracle as O
from tests.parity import compare_outputs, run_query


def test_scalar3d_gaussians(golden):
    g = golden("scalar3d_gaussians")
    t = O.tricubic(g["field"].copy(), "quiet")
    res, qa = run_query(t, g["query"])
    compare_outputs(res, qa, g, "", t, g["query"], tol=1e-12)
    assert np.array_equal(t.queryInds, g["queryInds"])


@pytest.mark.parametrize("mode", ["vector", "norm", "both", "bogus", "default"])
def test_dipole3d_modes(golden, mode):
    g = golden("dipole3d_modes")
    kw = {} if mode == "default" else {"mode": mode}
    t = O.tricubic(g["field"].copy(), "quiet", **kw)
    res, qa = run_query(t, g["query"])
    compare_outputs(res, qa, g, mode + "_", t, g["query"], tol=1e-12)
    assert np.array_equal(t.queryInds, g[mode + "_queryInds"])


def test_moving4d_scalar(golden):
    g = golden("moving4d_scalar")
    t = O.quadcubic(g["field"].copy(), "quiet")
    res, qa = run_query(t, g["query"])
    compare_outputs(res, qa, g, "", t, g["query"], tol=1e-12)
    assert np.array_equal(t.queryInds, g["queryInds"])


@pytest.mark.parametrize("mode", ["vector", "norm", "both", "default"])
def test_vector4d_modes(golden, mode):
    g = golden("vector4d_modes")
    kw = {} if mode == "default" else {"mode": mode}
    t = O.quadcubic(g["field"].copy(), "quiet", **kw)
    res, qa = run_query(t, g["query"])
    compare_outputs(res, qa, g, mode + "_", t, g["query"], tol=1e-12)
    assert np.array_equal(t.queryInds, g[mode + "_queryInds"])


def test_coefficient_tables(golden):
    g = golden("coeff_tables")
    t = O.tricubic(g["field3"].copy(), "quiet", mode="both")
    t.allCoeffs()
    assert np.array_equal(t.basePointInds, g["t3_basePointInds"])
    c = g["t3_consts"]
    p = t.p
    assert list(c[:9]) == list(p.h) + list(p.int_min) + list(p.int_max) and c[9] == t.nc
    assert np.array_equal(t.inputfield, g["t3_sorted"])
    for k, name in zip("xyzn", ("alphax", "alphay", "alphaz", "alphan")):
        ref = g["t3_" + name]
        got = t.tables[k]
        assert np.array_equal(np.isnan(got), np.isnan(ref))
        scale = np.abs(O.make_A(3)) @ np.abs(t.interps[k][t.p.base[None, :] +
                                                        O.neighbour_offsets(t.p.n, 3)[:, None]])
        m = ~np.isnan(ref[:, :-1])
        assert (np.abs(got[:, :-1] - ref[:, :-1])[m] <= 1e-15 * scale[m] + 1e-300).all()
    t4 = O.quadcubic(g["field4"].copy(), "quiet")
    t4.allCoeffs()
    assert np.array_equal(t4.basePointInds, g["t4_basePointInds"])
    assert np.array_equal(t4.inputfield, g["t4_sorted"])
    ref = g["t4_alphan"]
    got = t4.tables["n"]
    scale = np.abs(O.make_A(4)) @ np.abs(t4.interps["n"][t4.p.base[None, :] +
                                                         O.neighbour_offsets(t4.p.n, 4)[:, None]])
    assert (np.abs(got[:, :-1] - ref[:, :-1]) <= 1e-15 * scale).all()


def test_edge_battery(golden):
    g = golden("edge_battery")
    vrows = None
    from arbinterp_amd import workloads as W
    rows, _ = W.scalar3d_gaussians(seed=1)
    vrows, _ = W.dipole3d(seed=2, n=(14, 12, 10), nq=10)
    for i in range(int(g["ncases"])):
        pre = "c%03d_" % i
        fld = rows if bool(g[pre + "scalar"]) else vrows
        mode = str(g[pre + "mode"])
        kw = {} if mode == "none" else {"mode": mode}
        t = O.tricubic(fld.copy(), "quiet", **kw)
        res, qa = run_query(t, g[pre + "batch"])
        compare_outputs(res, qa, g, pre + "b_", t, g[pre + "batch"], tol=1e-12)
        t2 = O.tricubic(fld.copy(), "quiet", **kw)
        p = g[pre + "batch"][0]
        res1, qa1 = run_query(t2, p)
        compare_outputs(res1, qa1, g, pre + "s_", t2, p[None, :], tol=1e-12)


@pytest.mark.parametrize("tag", ["f32f32", "f64f32", "f32f64"])
def test_dtype_promotion(golden, tag):
    g = golden("dtype_promotion")
    t = O.tricubic(g[tag + "_field"].copy(), "quiet")
    res, qa = run_query(t, g[tag + "_query"])
    compare_outputs(res, qa, g, tag + "_", t, g[tag + "_query"], tol=1e-12)
    assert np.array_equal(t.queryInds, g[tag + "_queryInds"])

"""Write golden input/output vectors of the reference ARBInterp v1.8.

Run in a container where the reference is mounted (read-only, never copied):

    PYTHONDONTWRITEBYTECODE=1 python tests/golden/make_golden.py

It imports `tricubic` / `quadcubic` from the reference's `src/` tree and stores
inputs and outputs as `.npz` fixtures beside this script.  The fixtures are data
only; the reference's source never leaves its mount.  They pin what closed forms
and invariants leave open: end-to-end outputs at rounding level, the q == IntMax
aliasing outcomes per mode, the IndexError -> sQuery fallback, dtype promotion
of FP32 inputs, and the 4D row-shift quirk on general fields.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.abspath(os.path.join(HERE, "..", "..")))
REF_SRC = "/root/reference/src"

from arbinterp_amd import workloads as W  # noqa: E402


def _ref():
    sys.dont_write_bytecode = True
    sys.path.insert(0, REF_SRC)
    from ARBInterp.ARBInterp import tricubic, quadcubic  # noqa: E402
    return tricubic, quadcubic


def run_query(obj, q):
    """Call obj.Query on a copy; return (outputs tuple or exception, mutated q)."""
    qq = q.copy()
    try:
        out = obj.Query(qq)
    except Exception as e:  # noqa: BLE001 - the exception type is the datum
        return ("exc", type(e).__name__), qq
    if not isinstance(out, tuple):
        out = (out,)
    return ("ok", tuple(np.asarray(o, dtype=np.float64) for o in out)), qq


def pack(prefix, res, qafter, store):
    kind, val = res
    if kind == "exc":
        store[prefix + "exc"] = np.array(val)
    else:
        for i, o in enumerate(val):
            store[prefix + "out%d" % i] = o
    store[prefix + "qafter"] = qafter


def main():
    tricubic, quadcubic = _ref()
    os.makedirs(HERE, exist_ok=True)

    # 1. small scalar workload, full box queries
    rows, q = W.scalar3d_gaussians(seed=1)
    t = tricubic(rows.copy(), "quiet")
    res, qa = run_query(t, q)
    st = {"field": rows, "query": q}
    pack("", res, qa, st)
    st["queryInds"] = t.queryInds
    np.savez_compressed(os.path.join(HERE, "scalar3d_gaussians.npz"), **st)

    # 2. vector field, every mode; interior queries plus some outside the box
    rows, q = W.dipole3d(seed=2, n=(14, 12, 10), nq=1500)
    rng = np.random.default_rng(20)
    ext = np.array([13 * 1.0, 11 * 0.8, 9 * 1.25])
    qout = rng.uniform(-0.05, 1.05, size=(300, 3)) * ext
    q = np.concatenate([q, qout], axis=0)
    st = {"field": rows, "query": q}
    for mode in ["vector", "norm", "both", "bogus", None]:
        kw = {} if mode is None else {"mode": mode}
        t = tricubic(rows.copy(), "quiet", **kw)
        res, qa = run_query(t, q)
        key = "default" if mode is None else mode
        pack(key + "_", res, qa, st)
        st[key + "_queryInds"] = t.queryInds
    np.savez_compressed(os.path.join(HERE, "dipole3d_modes.npz"), **st)

    # 3. 4D scalar (moving Gaussians), small grid
    rows, q = W.moving4d(seed=3, n=(9, 8, 7, 6), h=(0.125, 0.125, 0.125, 0.2), nq=1200)
    rng = np.random.default_rng(30)
    lo = np.zeros(4)
    hi = np.array([8 * 0.125, 7 * 0.125, 6 * 0.125, 5 * 0.2])
    qout = rng.uniform(-0.05, 1.05, size=(200, 4)) * (hi - lo) + lo
    q = np.concatenate([q, qout], axis=0)
    t = quadcubic(rows.copy(), "quiet")
    res, qa = run_query(t, q)
    st = {"field": rows, "query": q}
    pack("", res, qa, st)
    st["queryInds"] = t.queryInds
    np.savez_compressed(os.path.join(HERE, "moving4d_scalar.npz"), **st)

    # 4. 4D vector field in every mode
    rng = np.random.default_rng(40)
    n4 = (7, 6, 6, 5)
    h4 = (0.5, 0.4, 0.45, 0.3)
    axes = [np.arange(n4[a]) * h4[a] + 0.25 * a for a in range(4)]
    X, Y, Z, T = np.meshgrid(*axes, indexing="ij")
    pos = np.array([[9.0, -4.0, 5.0], [-5.0, 8.0, -3.0]])
    mom = np.array([[1.0, 2.0, -0.5], [-1.5, 0.3, 2.0]])
    Bx, By, Bz = W.dipole_field_exact(pos, mom, X, Y, Z)
    s = 1.0 + 0.3 * np.sin(1.3 * T)
    rows = W.grid_rows(axes, [Bx * s, By * s, Bz * s], rng)
    lo = np.array([axes[a][1] for a in range(4)])
    hi = np.array([axes[a][-2] for a in range(4)])
    q = rng.uniform(size=(600, 4)) * (hi - lo) + lo
    st = {"field": rows, "query": q}
    for mode in ["vector", "norm", "both", None]:
        kw = {} if mode is None else {"mode": mode}
        t = quadcubic(rows.copy(), "quiet", **kw)
        res, qa = run_query(t, q)
        key = "default" if mode is None else mode
        pack(key + "_", res, qa, st)
        st[key + "_queryInds"] = t.queryInds
    np.savez_compressed(os.path.join(HERE, "vector4d_modes.npz"), **st)

    # 5. coefficient tables after allCoeffs on tiny grids
    st = {}
    rng = np.random.default_rng(50)
    n3 = (7, 6, 5)
    axes = [np.arange(n3[a]) * (0.3 + 0.1 * a) - 0.2 for a in range(3)]
    vals = [rng.normal(size=n3) for _ in range(3)]
    rows = W.grid_rows(axes, vals, rng)
    t = tricubic(rows.copy(), "quiet", mode="both")
    t.allCoeffs()
    st["field3"] = rows
    for k in ("alphax", "alphay", "alphaz", "alphan"):
        st["t3_" + k] = getattr(t, k)
    st["t3_basePointInds"] = t.basePointInds
    st["t3_consts"] = np.array([t.hx, t.hy, t.hz, t.xIntMin, t.yIntMin, t.zIntMin,
                                t.xIntMax, t.yIntMax, t.zIntMax, t.nc])
    st["t3_sorted"] = t.inputfield
    n4 = (6, 5, 5, 5)
    axes = [np.arange(n4[a]) * (0.2 + 0.05 * a) + 0.1 for a in range(4)]
    v4 = [rng.normal(size=n4)]
    rows = W.grid_rows(axes, v4, rng)
    t = quadcubic(rows.copy(), "quiet")
    t.allCoeffs()
    st["field4"] = rows
    st["t4_alphan"] = t.alphan
    st["t4_basePointInds"] = t.basePointInds
    st["t4_consts"] = np.array([t.hx, t.hy, t.hz, t.ht, t.xIntMin, t.yIntMin, t.zIntMin,
                                t.tIntMin, t.xIntMax, t.yIntMax, t.zIntMax, t.tIntMax, t.nc])
    st["t4_sorted"] = t.inputfield
    np.savez_compressed(os.path.join(HERE, "coeff_tables.npz"), **st)

    # 6. edge battery: exact IntMax / IntMin / node / NaN / outside queries
    edge = {}
    rows, _ = W.scalar3d_gaussians(seed=1)
    vrows, _ = W.dipole3d(seed=2, n=(14, 12, 10), nq=10)
    cases = []
    for fld, modes in ((rows, [None]), (vrows, ["vector", "norm", "both", None])):
        for mode in modes:
            kw = {} if mode is None else {"mode": mode}
            t = tricubic(fld.copy(), "quiet", **kw)
            lo = np.array([t.xIntMin, t.yIntMin, t.zIntMin])
            hi = np.array([t.xIntMax, t.yIntMax, t.zIntMax])
            mid = (lo + hi) / 2
            pts = []
            for a in range(3):
                for other in ("lo", "mid", "hi"):
                    p = {"lo": lo.copy(), "mid": mid.copy(), "hi": hi.copy()}[other]
                    p[a] = hi[a]
                    pts.append(p)
                    p2 = p.copy()
                    p2[a] = lo[a]
                    pts.append(p2)
            pts.append(hi.copy())
            pts.append(lo.copy())
            pts.append(np.array([np.nan, mid[1], mid[2]]))
            pts.append(np.array([lo[0] - 1e-9, mid[1], mid[2]]))
            pts.append(np.array([mid[0], hi[1] + 1e-9, mid[2]]))
            # a node in the middle of the grid
            pts.append(np.array([lo[0] + 3 * t.hx, lo[1] + 2 * t.hy, lo[2] + 1 * t.hz]))
            for p in pts:
                cases.append((fld is rows, mode, p))
    for i, (is_scalar, mode, p) in enumerate(cases):
        fld = rows if is_scalar else vrows
        kw = {} if mode is None else {"mode": mode}
        t = tricubic(fld.copy(), "quiet", **kw)
        batch = np.stack([p, np.array([t.xIntMin, t.yIntMin, t.zIntMin]) + 0.37 *
                          np.array([t.hx, t.hy, t.hz])], axis=0)
        res, qa = run_query(t, batch)
        pre = "c%03d_" % i
        edge[pre + "scalar"] = np.array(is_scalar)
        edge[pre + "mode"] = np.array("none" if mode is None else mode)
        edge[pre + "batch"] = batch
        pack(pre + "b_", res, qa, edge)
        t2 = tricubic(fld.copy(), "quiet", **kw)
        res1, qa1 = run_query(t2, p.copy())
        pack(pre + "s_", res1, qa1, edge)
    edge["ncases"] = np.array(len(cases))
    np.savez_compressed(os.path.join(HERE, "edge_battery.npz"), **edge)

    # 7. dtype promotion: FP32 field and/or FP32 queries
    st = {}
    rows, q = W.scalar3d_gaussians(seed=7, n=(12, 11, 10), nq=800)
    for tag, fdt, qdt in (("f32f32", np.float32, np.float32),
                          ("f64f32", np.float64, np.float32),
                          ("f32f64", np.float32, np.float64)):
        fr = rows.astype(fdt)
        qq = q.astype(qdt)
        t = tricubic(fr.copy(), "quiet")
        res, qa = run_query(t, qq)
        st[tag + "_field"] = fr
        st[tag + "_query"] = qq
        pack(tag + "_", res, qa, st)
        st[tag + "_queryInds"] = t.queryInds
    np.savez_compressed(os.path.join(HERE, "dtype_promotion.npz"), **st)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()

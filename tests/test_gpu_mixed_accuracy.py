"""Accuracy of every mixed-precision result the benchmark reports, against
the FP64 oracle.

The bound: 1e-6 relative error, normwise per query and output kind
(the 3 field components, the value, the d gradient components):

    ||o_q - r_q||_2 / max(||r_q||_2, 1e-4 ||S_q||_2)

with S the oracle's condition scale (the magnitude of the terms the
interpolant sums; oracle.rel_err_normwise).  Elementwise errors under the
same floor are printed beside it.  Workloads:

  * the headline: 128^3 dipole field, mode 'both', 10^6 interior queries on
    the default mixed path (node tiles, bucketed), every row checked;
  * moving4d: 48^3 x 16 scalar quadcubic, 2*10^6 queries (FP32 table);
  * the 64-field batch (FP32 tables, one persistent launch);
  * the particle push: 10^6 particles, 1000 RK4 steps of dt = 1e-3 in field
    0's well, mixed against FP64 final positions.
"""
import numpy as np
import pytest

from arbinterp_amd import workloads as W
from oracle import arbinterp_oracle as O

pytestmark = pytest.mark.gpu
TOLMIX = 1e-6


@pytest.fixture(scope="module")
def mods():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a ROCm GPU"
    from arbinterp_amd import ARBInterp as G
    from arbinterp_amd import batch as B
    return torch, G, B


def _errs(got, ref, scale, names):
    out = {}
    for name, a, r, s in zip(names, got, ref, scale):
        m = len(r)
        a = np.asarray(a, dtype=np.float64).reshape(m, -1)
        r = np.asarray(r, dtype=np.float64).reshape(m, -1)
        s = np.asarray(s, dtype=np.float64).reshape(m, -1)
        assert np.array_equal(np.isnan(a), np.isnan(r)), name
        nmx, nrms = O.rel_err_normwise(a, r, s)
        emx, _ = O.rel_err(a, r, s)
        out[name] = (nmx, nrms, emx)
        print("%-6s normwise max %.3e rms %.3e | elementwise max %.3e" % (name, nmx, nrms, emx))
    return out


def test_headline_mixed_accuracy(mods):
    torch, G, B = mods
    rows, q = W.dipole3d(seed=2, n=(128, 128, 128), nq=1_000_000)
    t = G.tricubic(rows, "quiet", mode="both", precision="mixed")
    qd = torch.from_numpy(q).cuda()
    c, nv, g = t.query_device(qd, out_f64=True)
    assert t.field_info().tiles_bytes > 0          # the node-tile path ran
    got = [x.cpu().numpy() for x in (c, nv, g)]
    t.close()
    o = O.tricubic(rows, "quiet", mode="both")
    ref = o.Query(q.copy())
    scale = o.condition_scale(q)
    e = _errs(got, ref, scale, ("comps", "norm", "grad"))
    for name, (nmx, _, _) in e.items():
        assert nmx <= TOLMIX, (name, nmx)


def test_moving4d_mixed_accuracy(mods):
    torch, G, B = mods
    rows, q = W.moving4d(seed=3)
    t = G.quadcubic(rows, "quiet", precision="mixed")
    qd = torch.from_numpy(q).cuda()
    _, nv, g = t.query_device(qd, out_f64=True)
    idx = np.random.default_rng(8).choice(len(q), 50_000, replace=False)
    got = [nv.cpu().numpy()[idx], g.cpu().numpy()[idx]]
    t.close()
    o = O.quadcubic(rows, "quiet")
    ref = o.Query(q[idx].copy())
    scale = o.condition_scale(q[idx])
    e = _errs(got, ref, scale, ("value", "grad"))
    for name, (nmx, _, _) in e.items():
        assert nmx <= TOLMIX, (name, nmx)


def test_multifield_batch_mixed_accuracy(mods):
    torch, G, B = mods
    fields = W.multi_field(seed=5)
    interps = [G.tricubic(r, "quiet", precision="mixed") for r, _ in fields]
    counts = [len(q) for _, q in fields]
    qd = torch.from_numpy(np.concatenate([q for _, q in fields])).cuda()
    bt = B.FieldBatch(interps)
    _, nv, g = bt.query(qd, counts, out_f64=True)
    nv = nv.cpu().numpy()
    g = g.cpu().numpy()
    bt.close()
    off = np.concatenate([[0], np.cumsum(counts)])
    rng = np.random.default_rng(12)
    worst = {"value": 0.0, "grad": 0.0}
    for k, (rows, q) in enumerate(fields):
        idx = rng.choice(len(q), min(len(q), 600), replace=False)
        o = O.tricubic(rows, "quiet")
        ref = o.Query(q[idx].copy())
        scale = o.condition_scale(q[idx])
        got = [nv[off[k] + idx], g[off[k] + idx]]
        for name, a, r, s in zip(("value", "grad"), got, ref, scale):
            m = len(idx)
            nmx, _ = O.rel_err_normwise(a.reshape(m, -1), r.reshape(m, -1), s.reshape(m, -1))
            worst[name] = max(worst[name], nmx)
        interps[k].close()
    print("64-field batch, mixed: normwise max value %.3e grad %.3e" % (worst["value"],
                                                                        worst["grad"]))
    assert worst["value"] <= TOLMIX and worst["grad"] <= TOLMIX, worst


def test_push_mixed_vs_fp64(mods):
    """The benchmark's push: 10^6 particles from rest in the central half of
    field 0's box (its bumps turned into wells), 1000 RK4 steps of 1e-3."""
    torch, G, B = mods
    rows = W.multi_field(seed=5)[0][0].copy()
    rows[:, 3] *= -1.0
    out = {}
    lost = {}
    for prec in ("fp64", "mixed"):
        t = G.tricubic(rows, "quiet", precision=prec)
        lo = np.array([t.xIntMin, t.yIntMin, t.zIntMin])
        hi = np.array([t.xIntMax, t.yIntMax, t.zIntMax])
        x0 = lo + (hi - lo) * (0.25 + 0.5 * np.random.default_rng(11).uniform(size=(1_000_000, 3)))
        x = torch.from_numpy(x0).cuda()
        v = torch.zeros_like(x)
        B.push_rk4(t, x, v, dt=1e-3, nsteps=1000, steps_per_launch=50, graph=True)
        torch.cuda.synchronize()
        out[prec] = x.cpu().numpy()
        lost[prec] = int(np.isnan(out[prec][:, 0]).sum())
        t.close()
    assert lost["mixed"] == lost["fp64"]
    m = ~np.isnan(out["fp64"]).any(axis=1)
    dx = np.abs(out["mixed"][m] - out["fp64"][m]).max()
    span = float(np.max(hi - lo))
    print("push: max |x_mixed - x_fp64| = %.3e (box %.3g), lost %d" % (dx, span, lost["fp64"]))
    assert dx <= 1e-5 * span

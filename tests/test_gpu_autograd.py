"""Autograd of the tricubic interpolant on the GPU.

Loss (random smooth scalar field, FP32 values):
    L = sum_i w_i U(x_i) + sum_i |grad U(x_i)|^2
Checked against (a) the oracle's FP64 reference gradient (Hessian-based for
positions, scatter-add adjoint for node values) and (b) FP64 central finite
differences (h = 1e-4 * dx) on sampled coordinates (kept away from cell faces:
the interpolant is only C1) and sampled nodes (exact: U is linear in them).
Bar: 1e-5 relative (per component, floored at 1e-3 of the largest)."""
import numpy as np
import pytest

from arbinterp_amd import workloads as W
from oracle import arbinterp_oracle as O

pytestmark = pytest.mark.gpu
FD_TOL = 1e-5


@pytest.fixture(scope="module")
def mods():
    import torch
    assert torch.cuda.is_available()
    from arbinterp_amd import ARBInterp as G
    from arbinterp_amd import autograd as AG
    return torch, G, AG


def _rel(a, b, floor=1e-3):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.maximum(np.abs(b), floor * np.abs(b).max())
    return float((np.abs(a - b) / den).max())


def _loss_terms(torch, U, G, w):
    return w * U + (G * G).sum(dim=1)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_backward_matches_oracle(mods, precision):
    torch, G, AG = mods
    rows, q, w = W.smooth3d(seed=4, n=(22, 20, 18), nq=20_000)
    t = G.tricubic(rows, "quiet", precision=precision)
    fld = AG.DifferentiableField(t, dtype=torch.float64)
    x = torch.from_numpy(q.astype(np.float64)).cuda().requires_grad_(True)
    wt = torch.from_numpy(w.astype(np.float64)).cuda()
    U, Gr = fld(x)
    L = (wt * U + (Gr * Gr).sum(dim=1)).sum()
    L.backward()
    o = O.tricubic(rows, "quiet")
    v, g = o.Query(q.astype(np.float64).copy())
    dq_ref, dval_ref = o.backward(q.astype(np.float64), w.astype(np.float64), 2 * g)
    # mixed: U's derivatives come from the node tiles (FP64 model, FP32
    # residuals).  On this coarse grid the field varies over a few cells, so
    # the residuals are of the order of the values, and this metric (floor
    # 1e-3 of the largest component, over 20 000 points) sees their FP32
    # rounding: measured 4.1e-5.  The finite-difference check at full size
    # holds the 1e-5 bar (measured 2.1e-6).
    tol = 1e-10 if precision == "fp64" else 5e-5
    e_q = _rel(x.grad.cpu().numpy(), dq_ref)
    e_v = _rel(fld.values.grad.cpu().numpy(), dval_ref)
    print("%s: d/dq rel %.2e  d/dvalues rel %.2e" % (precision, e_q, e_v))
    assert e_q < tol and e_v < tol


def test_backward_deterministic(mods):
    torch, G, AG = mods
    rows, q, w = W.smooth3d(seed=4, n=(30, 28, 26), nq=200_000)
    t = G.tricubic(rows, "quiet")
    fld = AG.DifferentiableField(t)
    grads = []
    for _ in range(2):
        x = torch.from_numpy(q.astype(np.float64)).cuda().requires_grad_(True)
        fld.values.grad = None
        U, Gr = fld(x)
        (U.sum() + (Gr * Gr).sum()).backward()
        grads.append((x.grad.clone(), fld.values.grad.clone()))
    assert torch.equal(grads[0][0], grads[1][0])
    assert torch.equal(grads[0][1], grads[1][1])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_backward_clustered(mods, precision):
    """Most queries in one cell (plus some outside the box): the node
    adjoint groups the queries by cell in index order with a stable radix
    sort, so a cell of tens of thousands of queries costs what it holds; the
    gradients match the oracle and repeat bit for bit."""
    torch, G, AG = mods
    rows, q, w = W.smooth3d(seed=6, n=(22, 20, 18), nq=12_000)
    t = G.tricubic(rows, "quiet", precision=precision)
    lo, hi = np.array(t._lo), np.array(t._hi)
    h = (hi - lo) / np.array(t.nPos[:3], dtype=np.float64)
    rng = np.random.default_rng(9)
    cell = lo + 0.41 * (hi - lo) + rng.uniform(0.1, 0.9, size=(48_000, 3)) * h
    out = lo - rng.uniform(0.5, 2.0, size=(2_000, 3)) * h
    qq = np.concatenate([cell, q, out])
    perm = rng.permutation(len(qq))
    qq = qq[perm]
    ww = np.concatenate([rng.normal(size=len(cell)), w, rng.normal(size=len(out))])[perm]
    fld = AG.DifferentiableField(t, dtype=torch.float64)
    grads = []
    for _ in range(2):
        x = torch.from_numpy(qq).cuda().requires_grad_(True)
        wt = torch.from_numpy(ww).cuda()
        fld.values.grad = None
        U, Gr = fld(x)
        ok = torch.isfinite(U)
        L = (wt[ok] * U[ok] + (Gr[ok] * Gr[ok]).sum(dim=1)).sum()
        L.backward()
        grads.append(fld.values.grad.clone())
    assert torch.equal(grads[0], grads[1])
    inside = np.isfinite(U.detach().cpu().numpy())
    assert inside.sum() == len(cell) + len(q)
    o = O.tricubic(rows, "quiet")
    v, g = o.Query(qq[inside].copy())
    _, dval_ref = o.backward(qq[inside], ww[inside], 2 * g)
    e_v = _rel(grads[0].cpu().numpy(), dval_ref)
    print("%s clustered: d/dvalues rel %.2e" % (precision, e_v))
    assert e_v < (1e-10 if precision == "fp64" else 5e-5)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_full_size_fd_checks(mods, precision):
    """96^3 field, 2^20 FP32 positions: the gradients of the FP64 or mixed
    path against central finite differences of the FP64 interpolant on 256
    coordinates and 256 nodes."""
    torch, G, AG = mods
    rows, q, w = W.smooth3d(seed=4)
    t = G.tricubic(rows, "quiet", precision=precision)
    fld = AG.DifferentiableField(t, dtype=torch.float32)
    x = torch.from_numpy(q).cuda().requires_grad_(True)          # FP32 positions
    wt = torch.from_numpy(w).cuda()
    U, Gr = fld(x)
    L = (wt * U + (Gr * Gr).sum(dim=1)).sum()
    L.backward()
    gq = x.grad.double().cpu().numpy()
    gv = fld.values.grad.double().cpu().numpy()
    # finite differences of the FP64 interpolant (a mixed forward's FP32
    # rounding, divided by 2e-4 h, would swamp them)
    t64 = t if precision == "fp64" else G.tricubic(rows, "quiet")
    vals64 = AG.node_values(t64)[0].clone()

    # positions: sample points whose cell-local coordinates are in [0.02, 0.98]
    q64 = q.astype(np.float64)
    loc = (q64 - np.array([t.xIntMin, t.yIntMin, t.zIntMin])) / np.array([t.hx, t.hy, t.hz])
    frac = loc - np.floor(loc)
    ok = np.where(((frac > 0.02) & (frac < 0.98)).all(axis=1))[0]
    idx = np.random.default_rng(1).choice(ok, 256, replace=False)
    delta = 1e-4 * t.hx
    pts = []
    for a in range(3):
        for sgn in (1.0, -1.0):
            p = q64[idx].copy()
            p[:, a] += sgn * delta
            pts.append(p)
    P = torch.from_numpy(np.concatenate(pts)).cuda()
    t64.allCoeffs()
    _, Up, Gp = t64.query_device(P, out_f64=True)
    Up = Up.cpu().numpy().reshape(6, 256)
    Gp = Gp.cpu().numpy().reshape(6, 256, 3)
    wi = w[idx].astype(np.float64)
    fd = np.zeros((256, 3))
    for a in range(3):
        lp = wi * Up[2 * a] + (Gp[2 * a] ** 2).sum(axis=1)
        lm = wi * Up[2 * a + 1] + (Gp[2 * a + 1] ** 2).sum(axis=1)
        fd[:, a] = (lp - lm) / (2 * delta)
    e_pos = _rel(gq[idx], fd)

    # nodes: perturb one node value at a time (FP64 node values)
    nodes = np.random.default_rng(2).choice(t._info.nnodes, 256, replace=False)
    qd = torch.from_numpy(q64).cuda()
    wd = torch.from_numpy(w.astype(np.float64)).cuda()
    eps = 1e-3
    fdn = np.zeros(256)
    for k, nidx in enumerate(nodes):
        terms = []
        for sgn in (1.0, -1.0):
            v = vals64.clone()
            v[nidx] += sgn * eps
            with torch.no_grad():
                Uv, Gv = AG.InterpFunction.apply(qd, v, t64)
            terms.append(wd * Uv + (Gv * Gv).sum(dim=1))
        fdn[k] = float((terms[0] - terms[1]).sum()) / (2 * eps)
    e_node = _rel(gv[nodes], fdn)
    print("FD (%s): positions rel %.2e, nodes rel %.2e" % (precision, e_pos, e_node))
    assert e_pos < FD_TOL and e_node < FD_TOL


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_backward_4d_matches_oracle(mods, precision):
    """Quadcubic field (moving Gaussians, 12 x 11 x 10 x 8 nodes): uniform
    positions, 3000 in one cell (more than one chunk of the per-cell sum) and
    500 outside the box.  d/dq and d/dvalues against the oracle's FP64
    adjoint; the node gradients repeat bit for bit."""
    torch, G, AG = mods
    rows, q = W.moving4d(seed=8, n=(12, 11, 10, 8), nq=6_000)
    t = G.quadcubic(rows.copy(), "quiet", precision=precision)
    lo, hi = np.array(t._lo), np.array(t._hi)
    h = (hi - lo) / np.array(t.nPos[:4], dtype=np.float64)
    rng = np.random.default_rng(12)
    cell = lo + 0.37 * (hi - lo) + rng.uniform(0.1, 0.9, size=(3_000, 4)) * h
    out = hi + rng.uniform(0.5, 2.0, size=(500, 4)) * h
    qq = np.concatenate([q, cell, out])
    qq = qq[rng.permutation(len(qq))]
    ww = rng.normal(size=len(qq))
    fld = AG.DifferentiableField(t, dtype=torch.float64)
    grads = []
    for _ in range(2):
        x = torch.from_numpy(qq).cuda().requires_grad_(True)
        wt = torch.from_numpy(ww).cuda()
        fld.values.grad = None
        U, Gr = fld(x)
        ok = torch.isfinite(U)
        L = (wt[ok] * U[ok] + (Gr[ok] * Gr[ok]).sum(dim=1)).sum()
        L.backward()
        grads.append((x.grad[ok].cpu().numpy(), fld.values.grad.clone()))
    assert torch.equal(grads[0][1], grads[1][1])
    inside = np.isfinite(U.detach().cpu().numpy())
    assert inside.sum() == len(q) + len(cell)
    o = O.quadcubic(rows.copy(), "quiet")
    v, g = o.Query(qq[inside].copy())
    dq_ref, dval_ref = o.backward(qq[inside], ww[inside], 2 * g)
    # mixed 4D reads the FP32 coefficient table: its rounding, seen through
    # this metric (floor 1e-3 of the largest component), bounds the error
    tol = 1e-10 if precision == "fp64" else 5e-5
    e_q = _rel(grads[0][0], dq_ref)
    e_v = _rel(grads[0][1].cpu().numpy(), dval_ref)
    print("4D %s: d/dq rel %.2e  d/dvalues rel %.2e" % (precision, e_q, e_v))
    assert e_q < tol and e_v < tol

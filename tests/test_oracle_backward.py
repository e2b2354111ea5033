"""The oracle's hessian() and backward(), pinned on the CPU.

They have no reference counterpart (the reference has no derivatives with
respect to its inputs), so they are pinned by what the mathematics fixes:

  * closed form: a field of degree <= 2 per axis is reproduced exactly by the
    interpolant (central-difference slopes are exact for quadratics), so its
    Hessian is the polynomial's own, constant Hessian;
  * central finite differences of the oracle's FP64 Query: the Hessian from
    differences of the gradient, d/dq of L = sum gv U + sum gg . grad U from
    differences of L (points at least 5 % of a cell from every face: the
    interpolant is only C1), and d/dvalues exactly (L is linear in the node
    values), in 3D and 4D.
"""
import numpy as np
import pytest

from arbinterp_amd import workloads as W
from oracle import arbinterp_oracle as O


def _grid(d, n, h):
    return W.grid_axes(n, h, origin=[0.3, -0.2, 0.1, 0.5][:d])


def _smooth(axes, seed):
    d = len(axes)
    rng = np.random.default_rng(seed)
    mesh = np.meshgrid(*axes, indexing="ij")
    ext = np.array([a[-1] - a[0] for a in axes])
    v = np.zeros(mesh[0].shape)
    for _ in range(3):
        c = np.array([a[0] for a in axes]) + rng.uniform(0.2, 0.8, size=d) * ext
        r2 = sum(((mesh[a] - c[a]) / (0.35 * ext[a])) ** 2 for a in range(d))
        v += rng.uniform(0.5, 2.0) * np.exp(-r2)
    return v


def _interior_points(o, n, rng):
    """Points whose local coordinates are in [0.05, 0.95] on every axis."""
    p = o.p
    d = o.NDIM
    cells = np.stack([rng.integers(0, p.npos[a], size=n) for a in range(d)], axis=1)
    u = rng.uniform(0.05, 0.95, size=(n, d))
    return np.stack([p.int_min[a] + (cells[:, a] + u[:, a]) * p.h[a] for a in range(d)], axis=1)


def test_hessian_of_quadratic_field_is_exact():
    axes = _grid(3, (9, 8, 10), (0.1, 0.13, 0.08))
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    H = np.array([[0.6, 0.2, 0.0], [0.2, -0.4, -0.1], [0.0, -0.1, 0.8]])
    f = (1.0 + 2 * X - Y + 0.5 * Z + 0.5 * (H[0, 0] * X * X + H[1, 1] * Y * Y + H[2, 2] * Z * Z)
         + H[0, 1] * X * Y + H[0, 2] * X * Z + H[1, 2] * Y * Z)
    o = O.tricubic(W.grid_rows(axes, [f]), "quiet")
    q = _interior_points(o, 500, np.random.default_rng(1))
    Ho = o.hessian(q)
    assert np.abs(Ho - H[None]).max() < 1e-10
    # and the gradient part of backward: gv * grad f with gg = 0
    gv = np.random.default_rng(2).normal(size=len(q))
    dq, _ = o.backward(q, gv, np.zeros((len(q), 3)))
    grad = np.stack([2 + H[0, 0] * q[:, 0] + H[0, 1] * q[:, 1] + H[0, 2] * q[:, 2],
                     -1 + H[0, 1] * q[:, 0] + H[1, 1] * q[:, 1] + H[1, 2] * q[:, 2],
                     0.5 + H[0, 2] * q[:, 0] + H[1, 2] * q[:, 1] + H[2, 2] * q[:, 2]], axis=1)
    assert np.abs(dq - gv[:, None] * grad).max() < 1e-9


def _loss(o, q, gv, gg):
    U, G = o.Query(q.copy())
    return gv * U[:, 0] + (gg * G).sum(axis=1)


@pytest.mark.parametrize("d", [3, 4])
def test_hessian_and_position_adjoint_by_finite_differences(d):
    n = (10, 9, 11) if d == 3 else (7, 6, 7, 6)
    h = (0.1, 0.12, 0.09, 0.15)[:d]
    axes = _grid(d, n, h)
    cls = O.tricubic if d == 3 else O.quadcubic
    o = cls(W.grid_rows(axes, [_smooth(axes, 3 + d)]), "quiet")
    rng = np.random.default_rng(4)
    q = _interior_points(o, 300, rng)
    gv = rng.normal(size=len(q))
    gg = rng.normal(size=(len(q), d))
    Ho = o.hessian(q)
    dq, _ = o.backward(q, gv, gg)
    eps = 1e-5 * min(h)
    Hfd = np.zeros_like(Ho)
    dqfd = np.zeros_like(dq)
    for a in range(d):
        qp, qm = q.copy(), q.copy()
        qp[:, a] += eps
        qm[:, a] -= eps
        Hfd[:, :, a] = (o.Query(qp.copy())[1] - o.Query(qm.copy())[1]) / (2 * eps)
        dqfd[:, a] = (_loss(o, qp, gv, gg) - _loss(o, qm, gv, gg)) / (2 * eps)
    sH = np.abs(Hfd).max()
    assert np.abs(Ho - Hfd).max() <= 1e-6 * sH
    assert np.abs(dq - dqfd).max() <= 1e-6 * np.abs(dqfd).max()


@pytest.mark.parametrize("d", [3, 4])
def test_node_adjoint_is_the_exact_linear_map(d):
    n = (8, 7, 9) if d == 3 else (6, 5, 6, 5)
    h = (0.1, 0.12, 0.09, 0.15)[:d]
    axes = _grid(d, n, h)
    vals = _smooth(axes, 7 + d)
    rows = W.grid_rows(axes, [vals])          # unshuffled: row = canonical node index
    cls = O.tricubic if d == 3 else O.quadcubic
    o = cls(rows.copy(), "quiet")
    rng = np.random.default_rng(5)
    q = _interior_points(o, 200, rng)
    gv = rng.normal(size=len(q))
    gg = rng.normal(size=(len(q), d))
    _, dval = o.backward(q, gv, gg)
    nodes = rng.choice(len(rows), 40, replace=False)
    delta = 1e-3
    for k in nodes:
        rp, rm = rows.copy(), rows.copy()
        rp[k, d] += delta
        rm[k, d] -= delta
        Lp = _loss(cls(rp, "quiet"), q, gv, gg).sum()
        Lm = _loss(cls(rm, "quiet"), q, gv, gg).sum()
        fd = (Lp - Lm) / (2 * delta)
        assert abs(dval[k] - fd) <= 1e-9 * max(1.0, np.abs(dval).max()), (k, dval[k], fd)

"""Pins of the CPU oracle by closed forms and invariants (no GPU)."""
import hashlib
import os

import numpy as np
import pytest

from oracle import arbinterp_oracle as O

# sha256 of the float64 matrices in the reference's examples/B_Matrix_3D.csv
# and B_Matrix_4D.csv (parsed with np.genfromtxt(..., delimiter=','))
B_SHA = {
    3: "095c8c49c2a3a83a2b9573c45b39ac968ad91054a34ddc62a0e05737c11aea65",
    4: "2c6cb7215ae0ff431a644ada06a741153efc9876bc5e315cb05309fe8a7389c5",
}
REF_EXAMPLES = "/root/reference/examples"


@pytest.mark.parametrize("d", [3, 4])
def test_B_matches_reference_csv(d):
    B = O.make_B(d)
    assert hashlib.sha256(np.ascontiguousarray(B).tobytes()).hexdigest() == B_SHA[d]
    path = os.path.join(REF_EXAMPLES, "B_Matrix_%dD.csv" % d)
    if os.path.exists(path):
        assert np.array_equal(B, np.genfromtxt(path, delimiter=","))


@pytest.mark.parametrize("d,den,amax", [(3, 8, 15.625), (4, 16, 39.0)])
def test_inverse_integer_and_A_dyadic(d, den, amax):
    Bi = np.linalg.inv(O.make_B(d))
    assert np.abs(Bi - np.round(Bi)).max() == 0.0
    A = O.make_A(d)
    assert np.abs(A * den - np.round(A * den)).max() == 0.0
    assert np.abs(A).max() == amax


def _grid_rows(n, h, origin, f, rng):
    axes = [origin[a] + np.arange(n[a]) * h[a] for a in range(len(n))]
    G = np.meshgrid(*axes, indexing="ij")
    vals = f(*G)
    d = len(n)
    cols = [g.transpose(*reversed(range(d))).reshape(-1) for g in G]
    v = vals.transpose(*reversed(range(d))).reshape(-1)
    rows = np.stack(cols + [v], axis=1)
    return rows[rng.permutation(len(rows))]


def test_triquadratic_reproduction():
    rng = np.random.default_rng(11)
    c = rng.normal(size=(3, 3, 3))
    f = lambda x, y, z: sum(c[a, b, e] * x ** a * y ** b * z ** e
                            for a in range(3) for b in range(3) for e in range(3))
    rows = _grid_rows((9, 8, 7), (0.1, 0.12, 0.15), (0.3, -0.2, 0.1), f, rng)
    t = O.tricubic(rows, "quiet")
    q = np.stack([rng.uniform(t.p.int_min[a], t.p.int_max[a], 400) for a in range(3)], 1)
    v, g = t.Query(q.copy())
    assert np.abs(v[:, 0] - f(*q.T)).max() < 1e-12
    # exact gradient of the polynomial
    fx = lambda x, y, z: sum(a * c[a, b, e] * x ** (a - 1) * y ** b * z ** e
                             for a in range(1, 3) for b in range(3) for e in range(3))
    assert np.abs(g[:, 0] - fx(*q.T)).max() < 1e-10


def test_4d_reproduction_without_xyzt_term():
    rng = np.random.default_rng(12)
    c = rng.normal(size=(3, 3, 3, 3))
    c[1:, 1:, 1:, 1:] = 0.0     # no term containing all four variables
    f = lambda x, y, z, t: sum(c[a, b, e, k] * x ** a * y ** b * z ** e * t ** k
                               for a in range(3) for b in range(3) for e in range(3)
                               for k in range(3))
    rows = _grid_rows((6, 6, 5, 5), (0.2, 0.25, 0.3, 0.4), (0.0, 0.1, 0.2, 0.3), f, rng)
    t = O.quadcubic(rows, "quiet")
    q = np.stack([rng.uniform(t.p.int_min[a], t.p.int_max[a], 200) for a in range(4)], 1)
    v, g = t.Query(q.copy())
    assert np.abs(v[:, 0] - f(*q.T)).max() < 1e-12


def test_4d_row_shift_known_answer():
    """f = xyzt: the reference's D leaves row 240 zero and shifts rows 241..255
    (ARBInterp.py:860), so the interpolant is f - hx*hy*hz*ht*H(u)H(v)H(w)H(tau)
    with H(s) = s(1-s)^2 (the Hermite basis function of d4f/dxdydzdt at corner 0)."""
    rng = np.random.default_rng(13)
    h = (0.3, 0.25, 0.2, 0.5)
    rows = _grid_rows((6, 5, 6, 5), h, (0.0, 0.1, 0.2, 0.3), lambda x, y, z, t: x * y * z * t, rng)
    t = O.quadcubic(rows, "quiet")
    q = np.stack([rng.uniform(t.p.int_min[a], t.p.int_max[a], 300) for a in range(4)], 1)
    v, _ = t.Query(q.copy())
    H = lambda s: s * (1 - s) ** 2
    loc = [((q[:, a] - t.p.int_min[a]) / t.p.h[a]) % 1.0 for a in range(4)]
    expect = q[:, 0] * q[:, 1] * q[:, 2] * q[:, 3] - np.prod(h) * np.prod([H(s) for s in loc], axis=0)
    assert np.abs(v[:, 0] - expect).max() < 1e-13
    # the quirk is visible: removing it would be off by up to ~prod(h)*(4/27)^4
    assert np.abs(np.prod(h) * np.prod([H(s) for s in loc], axis=0)).max() > 1e-6


def test_node_interpolation_is_central_difference():
    rng = np.random.default_rng(14)
    n, h = (8, 7, 6), (0.5, 0.4, 0.3)
    vals = rng.normal(size=n)
    axes = [np.arange(n[a]) * h[a] for a in range(3)]
    rows = _grid_rows(n, h, (0, 0, 0), lambda x, y, z: vals[np.rint(x / h[0]).astype(int),
                                                             np.rint(y / h[1]).astype(int),
                                                             np.rint(z / h[2]).astype(int)], rng)
    t = O.tricubic(rows, "quiet")
    pts, expv, expg = [], [], []
    for i in range(1, n[0] - 2):
        for j in range(1, n[1] - 2):
            for k in range(1, n[2] - 2):
                pts.append([axes[0][i], axes[1][j], axes[2][k]])
                expv.append(vals[i, j, k])
                expg.append([(vals[i + 1, j, k] - vals[i - 1, j, k]) / (2 * h[0]),
                             (vals[i, j + 1, k] - vals[i, j - 1, k]) / (2 * h[1]),
                             (vals[i, j, k + 1] - vals[i, j, k - 1]) / (2 * h[2])])
    v, g = t.Query(np.array(pts))
    assert np.abs(v[:, 0] - np.array(expv)).max() < 1e-13
    assert np.abs(g - np.array(expg)).max() < 1e-12


def _hermite(s):
    # 1D cubic Hermite basis: value at 0, value at 1, slope at 0, slope at 1
    return [2 * s ** 3 - 3 * s ** 2 + 1, -2 * s ** 3 + 3 * s ** 2,
            s ** 3 - 2 * s ** 2 + s, s ** 3 - s ** 2]


@pytest.mark.parametrize("d", [3, 4])
def test_independent_hermite_construction(d):
    """The interpolant equals the tensor-product cubic Hermite interpolant of
    the b-vector D.f (no inv(B) involved)."""
    rng = np.random.default_rng(15 + d)
    K, nco = 4 ** d, 2 ** d
    D = O.make_D(d)
    blocks = O.deriv_blocks(d)
    for _ in range(20):
        f = rng.normal(size=K)
        b = D @ f
        alpha = O.make_A(d) @ f
        s = rng.uniform(size=d)
        basis = [_hermite(s[a]) for a in range(d)]
        herm = 0.0
        for bi, blk in enumerate(blocks):
            for k in range(nco):
                w = 1.0
                for a in range(d):
                    corner = (k >> a) & 1
                    w *= basis[a][(2 if a in blk else 0) + corner]
                herm += b[bi * nco + k] * w
        mono = sum(alpha[i] * np.prod([s[a] ** ((i >> (2 * a)) & 3) for a in range(d)])
                   for i in range(K))
        assert abs(herm - mono) < 1e-12 * max(1.0, np.abs(b).sum())


def test_c1_continuity_across_faces():
    rng = np.random.default_rng(16)
    n, h = (9, 8, 7), (0.2, 0.3, 0.25)
    rows = _grid_rows(n, h, (0, 0, 0), lambda x, y, z: np.sin(3 * x) * np.cos(2 * y) + z * z, rng)
    t = O.tricubic(rows, "quiet")
    eps = 1e-9
    for a in range(3):
        for _ in range(30):
            p = np.array([rng.uniform(t.p.int_min[b] + 0.5 * h[b], t.p.int_max[b] - 0.5 * h[b])
                          for b in range(3)])
            k = rng.integers(2, n[a] - 3)
            p[a] = t.p.int_min[a] + k * h[a]
            lo, hi = p.copy(), p.copy()
            lo[a] -= eps
            hi[a] += eps
            v, g = t.Query(np.stack([lo, hi]))
            assert abs(v[0, 0] - v[1, 0]) < 1e-7
            assert np.abs(g[0] - g[1]).max() < 1e-6


def test_shuffle_invariance_and_base_points():
    rows, q = __import__("arbinterp_amd.workloads", fromlist=["x"]).scalar3d_gaussians(seed=1, nq=500)
    t1 = O.tricubic(rows, "quiet")
    t2 = O.tricubic(rows[::-1].copy(), "quiet")
    v1, g1 = t1.Query(q.copy())
    v2, g2 = t2.Query(q.copy())
    assert np.array_equal(np.nan_to_num(v1, nan=7.0), np.nan_to_num(v2, nan=7.0))
    assert np.array_equal(np.nan_to_num(g1, nan=7.0), np.nan_to_num(g2, nan=7.0))
    nx, ny, nz = t1.p.n
    base = [(i + 1) + nx * ((j + 1) + ny * (k + 1))
            for k in range(nz - 3) for j in range(ny - 3) for i in range(nx - 3)]
    assert np.array_equal(t1.basePointInds, np.array(base))


def test_rel_err_metric_floor():
    ref = np.array([1.0, 1e-9, -2.0])
    S = np.array([10.0, 10.0, 10.0])
    out = ref + np.array([1e-13, 1e-13, 0.0])
    mx, _ = O.rel_err(out, ref, S)
    assert abs(mx - 1e-13 / 1e-3) < 1e-20

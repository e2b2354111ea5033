"""Seeded synthetic fields and query sets for the five benchmark workloads.

Every workload draws from `np.random.default_rng` streams spawned from
`SeedSequence(seed)`: [field, shuffle, queries, extra].  Grid coordinates are
`origin + i*h` (exactly reproducible).  Fields are returned as the reference's
row format, an N x (d+1) or N x (d+3) array of (coords..., values...) with the
rows shuffled, which is what `tricubic(field)` / `quadcubic(field)` take
(README.md:12,17 of the reference).

  scalar3d_gaussians : 24x20x16 grid, h=(0.1,0.12,0.15), 4 Gaussians, 1e4
                       queries over the full box (edge points give NaN).
  dipole3d           : 128^3 grid, h=(1.0,0.8,1.25) mm, field of 6 point
                       magnetic dipoles outside the box, N x 7 rows, 1e7
                       interior queries, mode='both'.
  moving4d           : 48^3 x 16 grid, h=1/47 in space, dt=0.2, 3 moving
                       Gaussians with oscillating amplitude, 2e6 queries.
  smooth3d           : 96^3 grid, h=1, N(0,1) smoothed by a separable Gaussian
                       (sigma = 2 cells), FP32 values, 2^20 FP32 queries.
  multi_field        : 64 scalar fields, log-uniform edges in [16,160].
"""
from __future__ import annotations

import numpy as np

__all__ = ["streams", "grid_rows", "scalar3d_gaussians", "dipole3d", "moving4d",
           "smooth3d", "multi_field", "interior_queries", "dipole_field_exact"]


def streams(seed: int):
    return [np.random.default_rng(s) for s in np.random.SeedSequence(seed).spawn(4)]


def grid_axes(n, h, origin=None):
    origin = origin if origin is not None else [0.0] * len(n)
    return [origin[a] + np.arange(n[a]) * h[a] for a in range(len(n))]


def grid_rows(axes, values, shuffle_rng=None):
    """Rows (coords..., values...) in x-fastest order, optionally shuffled."""
    mesh = np.meshgrid(*axes, indexing="ij")
    d = len(axes)
    cols = [m.transpose(*reversed(range(d))).reshape(-1) for m in mesh]
    vals = [v.transpose(*reversed(range(d))).reshape(-1) for v in values]
    rows = np.stack(cols + vals, axis=1)
    if shuffle_rng is not None:
        rows = rows[shuffle_rng.permutation(len(rows))]
    return rows


def interior_queries(rng, lo, hi, n, dtype=np.float64):
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    q = rng.uniform(size=(n, len(lo))) * (hi - lo) + lo
    return q.astype(dtype)


def _gaussians(rng, k, lo, hi, wlo, whi):
    c = rng.uniform(lo, hi, size=(k, len(lo)))
    w = rng.uniform(wlo, whi, size=k)
    a = rng.uniform(0.5, 2.0, size=k)
    return c, w, a


def scalar3d_gaussians(seed=1, n=(24, 20, 16), h=(0.1, 0.12, 0.15), nq=10_000):
    """Scalar tricubic workload on a small grid: U = sum of 4 Gaussians; queries
    uniform over the full box so that edge points return NaN."""
    rf, rs, rq, _ = streams(seed)
    axes = grid_axes(n, h)
    lo = np.zeros(3)
    hi = np.array([(n[a] - 1) * h[a] for a in range(3)])
    c, w, amp = _gaussians(rf, 4, lo, hi, 0.3, 0.8)
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    U = np.zeros_like(X)
    for k in range(4):
        r2 = (X - c[k, 0]) ** 2 + (Y - c[k, 1]) ** 2 + (Z - c[k, 2]) ** 2
        U += amp[k] * np.exp(-r2 / (2 * w[k] ** 2))
    rows = grid_rows(axes, [U], rs)
    q = interior_queries(rq, lo, hi, nq)
    return rows, q


def _dipoles(rng, n, h):
    ext = np.array([(n[a] - 1) * h[a] for a in range(3)])
    centre = ext / 2
    half_diag = 0.5 * np.linalg.norm(ext)
    v = rng.normal(size=(6, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    rho = rng.uniform(1.2, 2.0, size=6) * half_diag
    pos = centre + rho[:, None] * v
    u = rng.normal(size=(6, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    mom = u * 10.0 ** rng.uniform(0.0, 1.0, size=(6, 1))
    return pos, mom


def dipole_field_exact(pos, mom, X, Y, Z):
    """B = sum_k (3 (m.rhat) rhat - m) / |r|^3 at the given points."""
    Bx = np.zeros_like(X)
    By = np.zeros_like(X)
    Bz = np.zeros_like(X)
    for p, m in zip(pos, mom):
        rx, ry, rz = X - p[0], Y - p[1], Z - p[2]
        r = np.sqrt(rx * rx + ry * ry + rz * rz)
        mr = (m[0] * rx + m[1] * ry + m[2] * rz) / r
        inv3 = 1.0 / r ** 3
        Bx += (3 * mr * rx / r - m[0]) * inv3
        By += (3 * mr * ry / r - m[1]) * inv3
        Bz += (3 * mr * rz / r - m[2]) * inv3
    return Bx, By, Bz


def dipole3d(seed=2, n=(128, 128, 128), h=(1.0, 0.8, 1.25), nq=10_000_000, shuffle=True):
    """Vector field of 6 point dipoles outside the box (N x 7 rows) and queries
    uniform in the interpolation interior [IntMin, IntMax]^3."""
    rf, rs, rq, _ = streams(seed)
    axes = grid_axes(n, h)
    pos, mom = _dipoles(rf, n, h)
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    Bx, By, Bz = dipole_field_exact(pos, mom, X, Y, Z)
    rows = grid_rows(axes, [Bx, By, Bz], rs if shuffle else None)
    lo = [axes[a][1] for a in range(3)]
    hi = [axes[a][-2] for a in range(3)]
    q = interior_queries(rq, lo, hi, nq)
    return rows, q


def moving4d(seed=3, n=(48, 48, 48, 16), h=None, nq=2_000_000, shuffle=True):
    """Scalar 4D field: 3 Gaussians moving linearly in t with amplitude
    1 + 0.3 sin(omega t)."""
    rf, rs, rq, _ = streams(seed)
    if h is None:
        h = (1.0 / 47, 1.0 / 47, 1.0 / 47, 0.2)
    axes = grid_axes(n, h)
    c = rf.uniform(0.0, 1.0, size=(3, 3))
    vel = rf.uniform(-0.1, 0.1, size=(3, 3))
    w = rf.uniform(0.1, 0.25, size=3)
    amp = rf.uniform(0.5, 2.0, size=3)
    om = rf.uniform(0.5, 2.0, size=3)
    X, Y, Z, T = np.meshgrid(*axes, indexing="ij")
    U = np.zeros_like(X)
    for k in range(3):
        r2 = ((X - c[k, 0] - vel[k, 0] * T) ** 2 + (Y - c[k, 1] - vel[k, 1] * T) ** 2
              + (Z - c[k, 2] - vel[k, 2] * T) ** 2)
        U += amp[k] * (1 + 0.3 * np.sin(om[k] * T)) * np.exp(-r2 / (2 * w[k] ** 2))
    rows = grid_rows(axes, [U], rs if shuffle else None)
    lo = [axes[a][1] for a in range(4)]
    hi = [axes[a][-2] for a in range(4)]
    q = interior_queries(rq, lo, hi, nq)
    return rows, q


def _smooth_axis(a, axis, sigma, truncate=4.0):
    r = int(truncate * sigma + 0.5)
    x = np.arange(-r, r + 1)
    k = np.exp(-0.5 * (x / sigma) ** 2)
    k /= k.sum()
    pad = [(0, 0)] * a.ndim
    pad[axis] = (r, r)
    ap = np.pad(a, pad, mode="symmetric")       # reflect about the edge
    out = np.zeros_like(a)
    n = a.shape[axis]
    for i, kv in enumerate(k):
        sl = [slice(None)] * a.ndim
        sl[axis] = slice(i, i + n)
        out += kv * ap[tuple(sl)]
    return out


def smooth3d(seed=4, n=(96, 96, 96), nq=1 << 20, shuffle=True):
    """Random smooth scalar field (FP32 values, FP64 coordinates, h = 1),
    FP32 query positions and N(0,1) loss weights."""
    rf, rs, rq, rx = streams(seed)
    axes = grid_axes(n, (1.0, 1.0, 1.0))
    U = rf.normal(size=n)
    for a in range(3):
        U = _smooth_axis(U, a, 2.0)
    U = U.astype(np.float32).astype(np.float64)
    rows = grid_rows(axes, [U], rs if shuffle else None)
    lo = [axes[a][1] for a in range(3)]
    hi = [axes[a][-2] for a in range(3)]
    q = interior_queries(rq, lo, hi, nq, dtype=np.float32)
    w = rx.normal(size=nq).astype(np.float32)
    return rows, q, w


def multi_field(seed=5, nfields=64, edge_lo=16, edge_hi=160, q_lo=1e3, q_hi=1e6, max_fields=None):
    """Many scalar fields of uneven size: each axis edge and each query count
    drawn log-uniformly; each field a 4-Gaussian mixture on the unit box."""
    rf, rs, rq, rx = streams(seed)
    out = []
    count = nfields if max_fields is None else min(nfields, max_fields)
    for i in range(count):
        n = tuple(int(round(np.exp(rx.uniform(np.log(edge_lo), np.log(edge_hi))))) for _ in range(3))
        nq = int(round(np.exp(rx.uniform(np.log(q_lo), np.log(q_hi)))))
        h = tuple(1.0 / (n[a] - 1) for a in range(3))
        axes = grid_axes(n, h)
        c, w, amp = _gaussians(rf, 4, np.zeros(3), np.ones(3), 0.1, 0.3)
        X, Y, Z = np.meshgrid(*axes, indexing="ij")
        U = np.zeros_like(X)
        for k in range(4):
            r2 = (X - c[k, 0]) ** 2 + (Y - c[k, 1]) ** 2 + (Z - c[k, 2]) ** 2
            U += amp[k] * np.exp(-r2 / (2 * w[k] ** 2))
        U = U.astype(np.float32).astype(np.float64)
        rows = grid_rows(axes, [U], rs)
        lo = [axes[a][1] for a in range(3)]
        hi = [axes[a][-2] for a in range(3)]
        q = interior_queries(rq, lo, hi, nq)
        out.append((rows, q))
    return out

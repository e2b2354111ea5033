"""arb_query_noalloc: the allocation-free variant of arb_query.

It runs the same kernels as arb_query with the bucketed path's sort
workspace carved from a fixed static device array, so its outputs must be
bitwise those of arb_query (comps, norms, grads, cell indices, NaN rows
written into the positions).  Checked through the C ABI on two calls in a
row, the second with fewer rows (the array keeps the first call's contents),
for the mixed 3D node-tile path, the FP64 3D table path, a mixed 4D field,
FP64 3D and 4D fields without a table (coefficients on the fly) and rows
clustered in one cell (buckets shared by several workgroups); device
memory must not change across the calls; bad sizes, refused
flags and unprepared fields give ARB_EINVAL.
"""
import ctypes

import numpy as np
import pytest

from arbinterp_amd import _native as N
from arbinterp_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a ROCm GPU"
    return torch


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream(torch):
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _outputs(torch, t, n, dt):
    info = t.field_info()
    d = t.NDIM
    comps = torch.full((n, 3), -7.0, dtype=dt, device="cuda") if info.has_comps else None
    norms = torch.full((n,), -7.0, dtype=dt, device="cuda") if info.has_norm else None
    grads = torch.full((n, d), -7.0, dtype=dt, device="cuda") if info.has_norm else None
    cells = torch.full((n,), -7, dtype=torch.int64, device="cuda")
    return comps, norms, grads, cells


def _call(torch, fn, t, q, outs, flags):
    comps, norms, grads, cells = outs
    return fn(t._h, _ptr(q), q.shape[0], q.stride(0), _ptr(comps), _ptr(norms), _ptr(grads),
              _ptr(cells), flags, _stream(torch))


def _same(a, b):
    if a is None:
        return b is None
    a = a.cpu().numpy()
    b = b.cpu().numpy()
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def _queries(t, nq, seed, clustered=False):
    """Positions over the box widened by a cell per side (some rows outside);
    clustered: four fifths of them inside one cell instead (a bucket of many
    more rows than one workgroup of the evaluation takes)."""
    rng = np.random.default_rng(seed)
    d = t.NDIM
    lo = np.array([getattr(t, _ax + "IntMin") for _ax in "xyzt"[:d]])
    hi = np.array([getattr(t, _ax + "IntMax") for _ax in "xyzt"[:d]])
    h = (hi - lo) / np.array(t.nPos[:d], dtype=np.float64)
    q = rng.uniform(lo - h, hi + h, size=(nq, d))
    if clustered:
        k = nq * 4 // 5
        q[:k] = lo + 0.37 * (hi - lo) + rng.uniform(0.1, 0.9, size=(k, d)) * h
        q = q[rng.permutation(nq)]
    return q


CASES = [
    # (name, field maker, constructor, kwargs, precision, flags, sizes)
    ("mixed3d_both_nodes", lambda: W.dipole3d(seed=7, n=(40, 36, 32), nq=1)[0], "tricubic",
     {"mode": "both"}, "mixed", N.ARB_Q_OUT_F64, (300_000, 120_000)),
    ("fp64_3d_both_table", lambda: W.dipole3d(seed=8, n=(40, 36, 32), nq=1)[0], "tricubic",
     {"mode": "both"}, "fp64", 0, (300_000, 120_000)),
    ("mixed3d_scalar_table", lambda: W.scalar3d_gaussians(seed=9, n=(36, 30, 28), nq=1)[0],
     "tricubic", {}, "mixed", N.ARB_Q_TABLE, (200_000, 70_000)),
    ("mixed4d_scalar", lambda: W.moving4d(seed=10, n=(20, 18, 16, 10), nq=1)[0], "quadcubic",
     {}, "mixed", 0, (250_000, 90_000)),
    # no table (a 1-byte table limit): coefficients formed on the fly
    ("fp64_3d_both_on_the_fly", lambda: W.dipole3d(seed=8, n=(40, 36, 32), nq=1)[0],
     "tricubic", {"mode": "both", "table_limit": 1}, "fp64", 0, (300_000, 120_000)),
    ("fp64_4d_scalar_on_the_fly", lambda: W.moving4d(seed=10, n=(20, 18, 16, 10), nq=1)[0],
     "quadcubic", {"table_limit": 1}, "fp64", 0, (250_000, 90_000)),
    # most rows in one cell: a bucket shared by many workgroups
    ("mixed3d_both_clustered", lambda: W.dipole3d(seed=7, n=(40, 36, 32), nq=1)[0], "tricubic",
     {"mode": "both"}, "mixed", 0, (300_000, 120_000), True),
    ("fp64_3d_both_clustered", lambda: W.dipole3d(seed=8, n=(40, 36, 32), nq=1)[0], "tricubic",
     {"mode": "both"}, "fp64", 0, (300_000, 120_000), True),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_noalloc_matches_query(torch, case):
    name, make, cls, kw, prec, flags, sizes = case[:7]
    clustered = len(case) > 7 and case[7]
    from arbinterp_amd import ARBInterp as G
    print("noalloc case", name, sizes, flush=True)
    t = getattr(G, cls)(make(), "quiet", precision=prec, **kw)
    L = N.lib()
    fl = flags | N.ARB_Q_SORTED | N.ARB_Q_NAN_ROWS
    dt = torch.float64 if (prec == "fp64" or flags & N.ARB_Q_OUT_F64) else torch.float32
    # prepare the field (table or node tiles) with an allocating call
    q0 = torch.zeros((1, t.NDIM), dtype=torch.float64, device="cuda")
    N.check(_call(torch, L.arb_query, t, q0[:0], _outputs(torch, t, 0, dt), fl), "prepare")
    otf = "table_limit" in kw
    if (prec == "fp64" or flags & N.ARB_Q_TABLE or t.NDIM == 4) and not otf:
        N.check(L.arb_build_coeffs(t._h, _stream(torch)), "build")
    assert t.field_info().on_the_fly == (1 if otf else 0)
    qs = [torch.from_numpy(_queries(t, n, 11 + k, clustered)).cuda()
          for k, n in enumerate(sizes)]
    refs, q_refs = [], []
    for q in qs:
        qr = q.clone()
        o = _outputs(torch, t, q.shape[0], dt)
        N.check(_call(torch, L.arb_query, t, qr, o, fl), "arb_query")
        refs.append(o)
        q_refs.append(qr)
    outs = [_outputs(torch, t, q.shape[0], dt) for q in qs]
    q_nas = [q.clone() for q in qs]
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for q, o in zip(q_nas, outs):
        N.check(_call(torch, L.arb_query_noalloc, t, q, o, fl), "arb_query_noalloc")
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    assert free1 == free0, "device memory changed by %d bytes" % (free0 - free1)
    for k in range(len(qs)):
        for a, b in zip(outs[k], refs[k]):
            assert _same(a, b), "%s call %d: outputs differ from arb_query" % (name, k)
        assert _same(q_nas[k], q_refs[k]), "%s call %d: NaN rows differ" % (name, k)
        nanrows = torch.isnan(q_nas[k]).any(dim=1).sum().item()
        assert 0 < nanrows < qs[k].shape[0], nanrows
    t.close()


def test_noalloc_refusals(torch):
    from arbinterp_amd import ARBInterp as G
    print("noalloc refusals", flush=True)
    L = N.lib()
    rows = W.dipole3d(seed=12, n=(24, 22, 20), nq=1)[0]
    t = G.tricubic(rows, "quiet", mode="both", precision="fp64")
    q = torch.from_numpy(_queries(t, 1000, 13)).cuda()
    o = _outputs(torch, t, 1000, torch.float64)
    # the table is not allocated yet
    assert _call(torch, L.arb_query_noalloc, t, q, o, 0) == N.ARB_EINVAL
    assert b"arb_build_coeffs" in L.arb_last_error()
    N.check(L.arb_build_coeffs(t._h, _stream(torch)), "build")
    assert _call(torch, L.arb_query_noalloc, t, q, o, 0) == N.ARB_OK
    # refused flags
    for bad in (N.ARB_Q_TIMING, N.ARB_Q_GROUP_ATOMIC):
        assert _call(torch, L.arb_query_noalloc, t, q, o, bad) == N.ARB_EINVAL
    # too many rows: refused before any work (the buffers hold every row anyway)
    nbig = (1 << 21) + 1
    qb = torch.from_numpy(_queries(t, nbig, 14)).cuda()
    ob = _outputs(torch, t, nbig, torch.float64)
    assert _call(torch, L.arb_query_noalloc, t, qb, ob, 0) == N.ARB_EINVAL
    assert b"limit" in L.arb_last_error()
    torch.cuda.synchronize()
    t.close()

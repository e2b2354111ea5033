"""Batched multi-field launch and the RK4 particle push on the GPU."""
import numpy as np
import pytest

from arbinterp_amd import workloads as W
from oracle import arbinterp_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods():
    import torch
    assert torch.cuda.is_available()
    from arbinterp_amd import ARBInterp as G
    from arbinterp_amd import batch as B
    return torch, G, B


def _eq(a, b):
    a = a.cpu().numpy()
    b = b.cpu().numpy()
    return np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(a[~np.isnan(a)],
                                                                       b[~np.isnan(b)])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_matches_single_launches(mods, precision):
    torch, G, B = mods
    from arbinterp_amd import _native as NN
    fields = W.multi_field(seed=5, nfields=8, edge_lo=12, edge_hi=40, q_lo=100, q_hi=20_000)
    interps = [G.tricubic(r, "quiet", precision=precision) for r, _ in fields]
    qs = [q for _, q in fields]
    rng = np.random.default_rng(3)
    qs = [np.concatenate([q, rng.uniform(-0.1, 1.1, size=(50, 3))]) for q in qs]
    counts = [len(q) for q in qs]
    qd = torch.from_numpy(np.concatenate(qs)).cuda()
    bt = B.FieldBatch(interps)
    bt.set_tuning(chunk=128, threads=128)
    _, nb, gb = bt.query(qd, counts)
    off = np.concatenate([[0], np.cumsum(counts)])
    for k, t in enumerate(interps):
        # the batch kernel reads the coefficient tables: the same bits as the
        # per-field table path (mixed 3D fields default to node tiles)
        _, n1, g1 = t.query_device(qd[off[k]:off[k + 1]].contiguous(),
                                   flags=NN.ARB_Q_DIRECT | NN.ARB_Q_TABLE)
        assert _eq(nb[off[k]:off[k + 1]], n1) and _eq(gb[off[k]:off[k + 1]], g1), k


def _rk4_oracle(o, x, v, dt, nsteps):
    def acc(p):
        _, g = o.Query(p.copy())
        return -g
    for _ in range(nsteps):
        a1 = acc(x)
        a2 = acc(x + 0.5 * dt * v)
        v2 = v + 0.5 * dt * a1
        a3 = acc(x + 0.5 * dt * v2)
        v3 = v + 0.5 * dt * a2
        a4 = acc(x + dt * v3)
        v4 = v + dt * a3
        x = x + dt / 6 * (v + 2 * v2 + 2 * v3 + v4)
        v = v + dt / 6 * (a1 + 2 * a2 + 2 * a3 + a4)
    return x, v


def test_rk4_graph_eager_and_oracle(mods):
    torch, G, B = mods
    rows, _ = W.multi_field(seed=5, nfields=1, edge_lo=40, edge_hi=41, q_lo=10, q_hi=11)[0]
    t = G.tricubic(rows, "quiet")
    rng = np.random.default_rng(9)
    n = 4096
    x0 = rng.uniform(0.25, 0.75, size=(n, 3))
    x0[:4] = [[0.06, 0.5, 0.5], [0.5, 0.95, 0.5], [0.5, 0.5, 0.03], [0.97, 0.97, 0.97]]
    v0 = rng.normal(scale=0.5, size=(n, 3))
    v0[:4] = [[-1.0, 0, 0], [0, 1.0, 0], [0, 0, -1.0], [1.0, 1.0, 1.0]]   # heading out of the box
    res = {}
    for graph in (False, True):
        x = torch.from_numpy(x0.copy()).cuda()
        v = torch.from_numpy(v0.copy()).cuda()
        B.push_rk4(t, x, v, dt=1e-3, nsteps=50, steps_per_launch=7, graph=graph)
        B.push_rk4(t, x, v, dt=1e-3, nsteps=50, steps_per_launch=7, graph=graph)
        torch.cuda.synchronize()
        res[graph] = (x.cpu().numpy(), v.cpu().numpy())
    for a, b in zip(res[False], res[True]):
        assert np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)])
    o = O.tricubic(rows, "quiet")
    sel = np.concatenate([np.arange(4), rng.choice(np.arange(4, n), 60, replace=False)])
    xo, vo = _rk4_oracle(o, x0[sel], v0[sel], 1e-3, 100)
    xg, vg = res[True][0][sel], res[True][1][sel]
    assert np.array_equal(np.isnan(xo), np.isnan(xg))
    m = ~np.isnan(xo)
    assert np.abs(xg[m] - xo[m]).max() < 1e-11
    assert np.isnan(xg[:4]).any()               # particles started near the box edge leave it


def test_rk4_mixed_vs_fp64(mods):
    torch, G, B = mods
    rows, _ = W.multi_field(seed=5, nfields=1, edge_lo=60, edge_hi=61, q_lo=10, q_hi=11)[0]
    out = {}
    rng = np.random.default_rng(4)
    x0 = rng.uniform(0.3, 0.7, size=(20_000, 3))
    for prec in ("fp64", "mixed"):
        t = G.tricubic(rows, "quiet", precision=prec)
        x = torch.from_numpy(x0.copy()).cuda()
        v = torch.zeros_like(x)
        B.push_rk4(t, x, v, dt=1e-3, nsteps=200)
        torch.cuda.synchronize()
        out[prec] = x.cpu().numpy()
    m = ~np.isnan(out["fp64"]).any(axis=1)
    dev = np.abs(out["mixed"][m] - out["fp64"][m]).max()
    print("mixed vs fp64 positions after 200 steps: max |dx| %.2e" % dev)
    assert np.array_equal(np.isnan(out["fp64"]), np.isnan(out["mixed"]))
    assert dev < 1e-5


def test_batch_follows_new_values_and_dead_fields(mods):
    """New node values (arb_field_set_values) reach the next batch query:
    doubling every node doubles every output bit for bit.  Destroying a
    field of the batch makes later batch queries raise instead of reading
    freed memory."""
    import ctypes
    torch, G, B = mods
    from arbinterp_amd import _native as NN
    fields = W.multi_field(seed=6, nfields=3, edge_lo=10, edge_hi=20, q_lo=500, q_hi=2000)
    interps = [G.tricubic(r, "quiet", precision="fp64") for r, _ in fields]
    counts = [len(q) for _, q in fields]
    qd = torch.from_numpy(np.concatenate([q for _, q in fields])).cuda()
    bt = B.FieldBatch(interps)
    _, n0, g0 = bt.query(qd, counts)
    n0, g0 = n0.clone(), g0.clone()
    t = interps[1]
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    v = torch.empty((t._info.nnodes,), dtype=torch.float64, device="cuda")
    NN.check(NN.lib().arb_field_values(t._h, ctypes.c_void_p(v.data_ptr()), st))
    v2 = 2.0 * v
    NN.check(NN.lib().arb_field_set_values(t._h, ctypes.c_void_p(v2.data_ptr()), st))
    _, n1, g1 = bt.query(qd, counts)
    a, b = counts[0], counts[0] + counts[1]
    assert _eq(n1[:a], n0[:a]) and _eq(n1[b:], n0[b:])
    assert _eq(n1[a:b], 2.0 * n0[a:b]) and _eq(g1[a:b], 2.0 * g0[a:b])
    interps[2].close()
    with pytest.raises(ValueError, match="destroyed"):
        bt.query(qd, counts)
    bt.close()
    for k in (0, 1):
        interps[k].close()


def test_out_buffers_are_checked(mods):
    """Caller-supplied output buffers of the wrong shape, dtype, device or
    layout are refused before any kernel runs."""
    torch, G, B = mods
    rows, q = W.dipole3d(seed=4, n=(14, 13, 12), nq=1000)
    t = G.tricubic(rows, "quiet", mode="both", precision="mixed")
    qd = torch.from_numpy(q).cuda()
    f32 = dict(dtype=torch.float32, device="cuda")
    good = (torch.empty((1000, 3), **f32), torch.empty((1000,), **f32),
            torch.empty((1000, 3), **f32))
    t.query_device(qd, out=good)
    bad_sets = [
        (torch.empty((999, 3), **f32), good[1], good[2]),                  # too few rows
        (good[0], good[1].double(), good[2]),                              # dtype
        (good[0], good[1], torch.empty((1000, 3), dtype=torch.float32)),   # host memory
        (good[0], good[1], torch.empty((3, 1000), **f32).t()),             # not contiguous
        (good[0], None, good[2]),                                          # missing
    ]
    for bad in bad_sets:
        with pytest.raises(ValueError):
            t.query_device(qd, out=bad)
    W8 = int(t.field_info().packed_width)
    with pytest.raises(ValueError):
        t.query_packed(qd, out=torch.empty((1000, W8), **f32), out_f64=True)
    base = torch.empty((1000 * W8 + 1,), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="aligned"):
        t.query_packed(qd, out=base[1:].view(1000, W8), out_f64=True)
    with pytest.raises(ValueError):
        t.query_device(qd.float())
    bt = B.FieldBatch([t])
    with pytest.raises(ValueError):
        bt.query(qd, [1001])
    with pytest.raises(ValueError):
        bt.query(qd, [1000], out=(good[0], good[1], torch.empty((1000, 4), **f32)))
    bt.close()
    t.close()

"""Grouping of the rows by bucket on the bucketed query path.

The bucketed kernels need the rows grouped by bucket, in any order inside a
bucket (each row is evaluated alone).  Up to 2^24 rows and 2^16 bucket keys
the rows are grouped by the high 8 bits of the key, then each of those 256
segments by its low bits, with no stable ranking; past either limit, and
with the cell in the key (ARB_Q_SORT_CELL), by the LSD radix sort.  Every
grouping gives the same bits as the direct path, which evaluates each row
where it stands (node tiles for mixed 3D fields, the coefficient table
otherwise).  Pinned here on the distributions that stress the grouping:
uniform rows; every row in one cell (one key: a segment of many tiles,
and a bucket shared by many evaluation workgroups, on every bucketed
kernel);
rows in a few neighbouring buckets; half the rows outside the box (the
sentinel key, the largest); row counts at tile boundaries; more than 2^24
rows (the radix sort), compared on the device; keys of 9, 16 and 17 bits
(either side of the grouping's limits).
"""
import numpy as np
import pytest

from arbinterp_amd import workloads as W

pytestmark = pytest.mark.gpu
TILE = 8192                      # rows per tile of the sort passes


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a ROCm GPU"
    from arbinterp_amd import ARBInterp as G
    from arbinterp_amd import _native as NN
    rows, _ = W.dipole3d(seed=5, n=(48, 40, 36), nq=1)
    fields = {p: G.tricubic(rows.copy(), "quiet", mode="both", precision=p)
              for p in ("mixed", "fp64")}
    lo, hi = np.array(fields["fp64"]._lo), np.array(fields["fp64"]._hi)
    yield torch, G, NN, fields, lo, hi
    for t in fields.values():
        t.close()


def _query_sets(lo, hi):
    rng = np.random.default_rng(21)
    span = hi - lo
    sets = {"uniform": rng.uniform(lo, hi, size=(150_000, 3))}
    c = lo + 0.37 * span
    h = span / 200.0                  # well inside one cell
    sets["one_cell"] = c + rng.uniform(0.0, 1.0, size=(100_000, 3)) * h
    sets["few_buckets"] = lo + 0.5 * span + rng.uniform(-0.04, 0.04, size=(120_000, 3)) * span
    inside = rng.uniform(lo, hi, size=(40_000, 3))
    outside = rng.uniform(lo - 0.5 * span, lo - 0.01 * span, size=(40_000, 3))
    half = np.concatenate([inside, outside])
    sets["half_outside"] = half[rng.permutation(len(half))]
    for n in (1, TILE - 1, TILE, TILE + 1, 3 * TILE + 17):
        sets["n%d" % n] = rng.uniform(lo, hi, size=(n, 3))
    return sets


def _run(torch, t, q, flags):
    qd = torch.from_numpy(np.ascontiguousarray(q)).cuda()
    out = t.query_device(qd, out_f64=True, flags=flags)
    torch.cuda.synchronize()
    return [x.cpu().numpy() if x is not None else None for x in out]


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_grouped_rows_give_direct_bits(env, prec):
    torch, G, NN, fields, lo, hi = env
    t = fields[prec]
    for name, q in _query_sets(lo, hi).items():
        ref = _run(torch, t, q, NN.ARB_Q_DIRECT)
        for fname, fl in (("grouped", NN.ARB_Q_SORTED),
                          ("radix", NN.ARB_Q_SORTED | NN.ARB_Q_SORT_CELL)):
            got = _run(torch, t, q, fl)
            for a, b in zip(ref, got):
                assert np.array_equal(a, b, equal_nan=True), (prec, name, fname)
        if name == "half_outside":
            assert np.isnan(ref[1]).sum() >= 30_000


def test_grouped_4d_gives_direct_bits(env):
    torch, G, NN, _, _, _ = env
    rows, q = W.moving4d(seed=4, n=(20, 18, 16, 10), nq=60_000)
    for prec in ("mixed", "fp64"):
        t = G.quadcubic(rows.copy(), "quiet", mode="norm", precision=prec)
        for qq in (q, np.tile(q[:1], (20_000, 1))):
            ref = _run(torch, t, qq, NN.ARB_Q_DIRECT)
            got = _run(torch, t, qq, NN.ARB_Q_SORTED)
            info = t.field_info()
            nbt = int(np.prod([-(-info.npos[a] // info.block[a]) for a in range(4)]))
            assert nbt >= 256, nbt                 # keys of 9 bits or more: grouped
            for a, b in zip(ref, got):
                if a is not None:
                    assert np.array_equal(a, b, equal_nan=True), prec
        t.close()


@pytest.mark.parametrize("d,prec,kind", [
    (3, "fp64", "table"), (3, "fp64", "on_the_fly"), (3, "mixed", "table"),
    (3, "mixed", "nodes"), (4, "fp64", "table"), (4, "mixed", "table"),
    (4, "fp64", "on_the_fly")])
def test_hot_bucket_every_kernel(env, d, prec, kind):
    """A bucket holding many more rows than one workgroup takes (every row
    of the set in one cell, plus a uniform spread) is shared by several
    workgroups on every bucketed kernel: staged table, on-the-fly (3D and
    4D) and node tiles; the bits are the direct path's on the table."""
    torch, G, NN, _, _, _ = env
    if d == 3:
        rows, _ = W.dipole3d(seed=9, n=(30, 26, 22), nq=1)
        cls, kw = G.tricubic, {"mode": "both"}
    else:
        rows, _ = W.moving4d(seed=6, n=(14, 12, 12, 8), nq=1)
        cls, kw = G.quadcubic, {"mode": "norm"}
    ref_t = cls(rows.copy(), "quiet", precision=prec, **kw)
    lo, hi = np.array(ref_t._lo), np.array(ref_t._hi)
    rng = np.random.default_rng(17)
    span = hi - lo
    cell = lo + 0.43 * span + rng.uniform(0.0, 1e-3, size=(30_000, d)) * span
    q = np.concatenate([cell, rng.uniform(lo, hi, size=(5_000, d))])
    q = q[rng.permutation(len(q))]
    base = NN.ARB_Q_TABLE if (d == 3 and prec == "mixed" and kind == "table") else 0
    ref = _run(torch, ref_t, q, base | NN.ARB_Q_DIRECT)
    t = cls(rows.copy(), "quiet", precision=prec, table_limit=1, **kw) \
        if kind == "on_the_fly" else ref_t
    got = _run(torch, t, q, base | NN.ARB_Q_SORTED)
    if kind == "on_the_fly":
        assert t.field_info().on_the_fly == 1
    for a, b in zip(ref, got):
        if a is not None:
            assert np.array_equal(a, b, equal_nan=True), (d, prec, kind)
    if t is not ref_t:
        t.close()
    ref_t.close()


def test_more_rows_than_the_grouping_takes(env):
    """2^24 + 5 rows (the radix sort) give the direct path's bits; so do
    2^24 - 1 rows (the grouping), compared on the device."""
    torch, G, NN, fields, lo, hi = env
    t = fields["mixed"]
    g = torch.Generator(device="cuda")
    g.manual_seed(8)
    lo_t = torch.tensor(lo, dtype=torch.float64, device="cuda")
    hi_t = torch.tensor(hi, dtype=torch.float64, device="cuda")
    for n in ((1 << 24) - 1, (1 << 24) + 5):
        q = lo_t + (hi_t - lo_t) * torch.rand((n, 3), dtype=torch.float64, device="cuda",
                                             generator=g)
        ref = t.query_device(q, flags=NN.ARB_Q_DIRECT)
        got = t.query_device(q, flags=NN.ARB_Q_SORTED)
        torch.cuda.synchronize()
        for a, b in zip(ref, got):
            assert torch.equal(a, b), n
        del q, ref, got
        torch.cuda.empty_cache()


@pytest.mark.parametrize("nodes,block,bits", [
    ((24, 22, 20), (4, 4, 2), 9),       # 6 x 5 x 9 = 270 buckets: the grouping's narrowest keys
    ((41, 41, 41), (1, 1, 1), 16),      # 38^3 = 54872 buckets: its widest
    ((44, 44, 44), (1, 1, 1), 17)])     # 41^3 = 68921 buckets: the radix sort
def test_key_widths_at_the_grouping_limits(env, nodes, block, bits):
    """Bucket keys of 9 and 16 bits (the grouping's limits) and 17 bits (the
    radix sort), forced by the cell block: the grouped rows give the direct
    path's bits, FP64 and mixed, on uniform rows and on rows in one cell."""
    torch, G, NN, _, _, _ = env
    rows, _ = W.dipole3d(seed=13, n=nodes, nq=1)
    for prec in ("mixed", "fp64"):
        t = G.tricubic(rows.copy(), "quiet", mode="both", precision=prec)
        t.set_tuning(block)
        info = t.field_info()
        assert [int(info.block[a]) for a in range(3)] == list(block)
        nbt = int(np.prod([-(-info.npos[a] // block[a]) for a in range(3)]))
        assert nbt.bit_length() == bits, (nbt, bits)     # keys 0..nbt
        lo, hi = np.array(t._lo), np.array(t._hi)
        rng = np.random.default_rng(bits)
        span = hi - lo
        sets = {"uniform": rng.uniform(lo, hi, size=(70_000, 3)),
                "one_cell": lo + 0.61 * span + rng.uniform(0.0, 1.0, size=(20_000, 3)) * span / 400}
        for name, q in sets.items():
            ref = _run(torch, t, q, NN.ARB_Q_DIRECT)
            got = _run(torch, t, q, NN.ARB_Q_SORTED)
            for a, b in zip(ref, got):
                if a is not None:
                    assert np.array_equal(a, b, equal_nan=True), (prec, name, bits)
        t.close()

"""Autotuner: candidates that must not depend on the launch shape give
bit-identical results, node-tile candidates agree with the direct kernel to
FP32 rounding (the tuner raises otherwise), and the chosen shape is
applied."""
import numpy as np
import pytest

from arbinterp_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_tune_query_and_batch():
    import torch
    assert torch.cuda.is_available()
    from arbinterp_amd.ARBInterp import tricubic
    from arbinterp_amd import autotune as AT
    from arbinterp_amd import batch as B
    rows, q = W.dipole3d(seed=2, n=(40, 36, 33), nq=200_000)
    t = tricubic(rows, "quiet", mode="both", precision="mixed")
    qd = torch.from_numpy(q).cuda()
    ref = [x.clone() for x in t.query_device(qd)]
    res = AT.tune_query(t, qd, reps=1, use_cache=False)
    assert res["path"] in ("direct", "table", "sorted") and len(res["table"]) > 3
    assert {c[0][0] for c in res["table"]} == {"direct", "table", "sorted"}
    again = t.query_device(qd)
    assert AT._close(again, ref)
    # the applied shape is what later default queries run: same bits again
    for a, b in zip(again, t.query_device(qd)):
        assert torch.equal(a, b)
    fields = W.multi_field(seed=5, nfields=4, edge_lo=12, edge_hi=30, q_lo=500, q_hi=5000)
    interps = [tricubic(r, "quiet") for r, _ in fields]
    counts = [len(qq) for _, qq in fields]
    qb = torch.from_numpy(np.concatenate([qq for _, qq in fields])).cuda()
    bt = B.FieldBatch(interps)
    rb = AT.tune_batch(bt, qb, counts, reps=1)
    assert rb["chunk"] in (64, 128, 256, 512, 1024)

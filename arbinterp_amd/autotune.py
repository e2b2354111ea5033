"""Autotuner for the query kernels.

`tune_query(interp, q)` times the candidate launch shapes of one interpolator
on the caller's own positions (GPU, HIP events on the current stream):

  * the direct lane-per-query kernel;
  * the cell-bucketed kernel for every block of cells (cells per axis) on a
    ladder whose LDS footprint fits, with 64/128/256 threads per workgroup.

Every candidate's outputs are compared bit for bit with the first one (the
result does not depend on the launch shape; a difference raises).  The
fastest shape is applied (`interp.set_tuning`, `interp.default_flags`) and
remembered per (grid, mode, precision, query-count octave) for the process.

`tune_batch(batch, q, counts)` does the same for the batched launch (queries
per work item, threads per workgroup, workgroups per compute unit).
"""
from __future__ import annotations

import itertools
import math

import torch

from . import _native as N

_CACHE: dict = {}


def _key(interp, n):
    info = interp._info
    return (interp.NDIM, tuple(int(info.n[a]) for a in range(interp.NDIM)), int(info.mode),
            int(info.precision), int(math.log2(max(1, n))))


def block_ladder(interp, max_lds=64 * 1024):
    """Candidate cell blocks (one edge per axis) whose LDS stage fits."""
    d = interp.NDIM
    K = 4 ** d
    esz = 8 if interp.precision == "fp64" else 4
    cellbytes = (interp._info.ninterp * K * esz // 16 + 1) * 16
    edges = (1, 2, 4, 8)
    out = []
    for b in itertools.product(edges, repeat=d):
        cells = math.prod(b)
        if cells * cellbytes > max_lds or cells > 128:
            continue
        if any(b[a] > int(interp._info.npos[a]) for a in range(d)):
            continue
        out.append(b)
    return out


def _same(a, b):
    for x, y in zip(a, b):
        if x is None and y is None:
            continue
        if not torch.equal(torch.nan_to_num(x, nan=1e300), torch.nan_to_num(y, nan=1e300)):
            return False
    return True


def _time(fn, reps):
    fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def tune_query(interp, q, reps=3, max_candidates=None, verify=True, use_cache=True):
    """Pick the fastest launch shape for queries like `q` (CUDA float64 tensor).
    Returns {"path", "block", "threads", "ms", "table": [(cand, ms), ...]}."""
    n = q.shape[0]
    key = _key(interp, n)
    if use_cache and key in _CACHE:
        best = _CACHE[key]
        _apply(interp, best)
        return dict(best, cached=True)
    out = interp.query_device(q, flags=N.ARB_Q_DIRECT)
    ref = tuple(t.clone() if t is not None else None for t in out)
    cands = [("direct", None, 0)]
    for b in block_ladder(interp):
        for th in (64, 128, 256):
            cands.append(("sorted", b, th))
    if max_candidates:
        cands = cands[:max_candidates]
    table = []
    for path, b, th in cands:
        if path == "direct":
            interp.set_tuning(None, 0)
            fl = N.ARB_Q_DIRECT
        else:
            interp.set_tuning(b, th)
            fl = N.ARB_Q_SORTED
        fn = lambda: interp.query_device(q, flags=fl, out=out)
        ms = _time(fn, reps)
        if verify and not _same(out, ref):
            raise RuntimeError("autotune: launch shape %s %s %d changed the results" % (path, b, th))
        table.append(((path, b, th), ms))
    (path, b, th), ms = min(table, key=lambda x: x[1])
    best = {"path": path, "block": b, "threads": th, "ms": ms}
    _apply(interp, best)
    _CACHE[key] = best
    return dict(best, table=table, cached=False)


def _apply(interp, best):
    if best["path"] == "direct":
        interp.set_tuning(None, 0)
        interp.default_flags = N.ARB_Q_DIRECT
    else:
        interp.set_tuning(best["block"], best["threads"])
        interp.default_flags = N.ARB_Q_SORTED


def tune_batch(batch, q, counts, reps=3, verify=True):
    """Pick (chunk, threads, blocks_per_cu) for a FieldBatch and these counts."""
    ref = tuple(t.clone() if t is not None else None for t in batch.query(q, counts))
    out = batch.query(q, counts)
    table = []
    for chunk, th, bpc in itertools.product((64, 128, 256, 512, 1024), (64, 128, 256), (2, 4, 8)):
        if th > chunk:
            continue
        batch.set_tuning(chunk, th, bpc)
        ms = _time(lambda: batch.query(q, counts, out=out), reps)
        if verify and not _same(out, ref):
            raise RuntimeError("autotune: batch shape (%d,%d,%d) changed the results" % (chunk, th, bpc))
        table.append(((chunk, th, bpc), ms))
    (chunk, th, bpc), ms = min(table, key=lambda x: x[1])
    batch.set_tuning(chunk, th, bpc)
    return {"chunk": chunk, "threads": th, "blocks_per_cu": bpc, "ms": ms, "table": table}

"""Benchmark: value+gradient queries/s of the tricubic interpolant on one GPU.

Workload: 128^3 vector field of 6 point dipoles (h = 1.0, 0.8, 1.25 mm, N x 7
rows), mode='both' (components, |B| and grad|B|), 10^7 query positions
uniform in the interpolation interior, already resident in HBM.  One step =
one batch query of all 10^7 positions.  The headline path is mixed precision
(FP64 cell reduction and FD stencils, FP32 coefficient table and evaluation);
the plain FP64 path is timed beside it.

Prints one JSON line (rank 0).  `--gpus N` with N > 1 is not supported: this
build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

for _v in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
    os.environ.setdefault(_v, "1")

import numpy as np  # noqa: E402

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "value+grad queries/s, mixed-prec vs FP64 plain on 1 GPU at ≤1e-6 rel err"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nq", type=int, default=10_000_000)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--cpu-chunks", type=int, default=6,
                    help="2^16-query chunks timed for the CPU baseline (0: skip)")
    ap.add_argument("--err-sample", type=int, default=200_000)
    ap.add_argument("--flags", type=int, default=0, help="extra ARB_Q_* flags")
    ap.add_argument("--variants", default="", help="comma list of query paths to time "
                    "(auto,direct,sorted_lds,sorted_global); default all")
    ap.add_argument("--timing-steps", type=int, default=3)
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_steps(torch, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(steps):
        fn()
    ev1.record()
    torch.cuda.synchronize()
    return ev0.elapsed_time(ev1) / steps


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only (got --gpus %d); this build has no multi-GPU path"
              % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    from arbinterp_amd import workloads as W
    from arbinterp_amd.ARBInterp import tricubic

    n = (args.grid,) * 3
    t0 = time.time()
    rows, q = W.dipole3d(seed=2, n=n, nq=args.nq)
    log("workload generated in %.1f s" % (time.time() - t0))
    qd = torch.from_numpy(q).cuda()

    from arbinterp_amd import _native as NN
    variants = {"auto": 0, "direct": NN.ARB_Q_DIRECT, "sorted_lds": NN.ARB_Q_SORTED,
                "sorted_global": NN.ARB_Q_SORTED_GLOBAL}
    if args.variants:
        variants = {k: v for k, v in variants.items() if k in args.variants.split(",")}
    res = {}
    for prec in ("mixed", "fp64"):
        t = tricubic(rows, "quiet", mode="both", precision=prec)
        torch.cuda.synchronize()
        tb = time.time()
        t.allCoeffs()
        build_ms = (time.time() - tb) * 1e3
        out = t.query_device(qd, cells=True)
        torch.cuda.synchronize()
        per = {}
        for vname, fl in variants.items():
            fn = lambda: t.query_device(qd, out=out[:3], flags=fl | args.flags)
            ms = time_steps(torch, fn, args.steps, args.warmup)
            st = np.zeros(4)
            for _ in range(args.timing_steps):
                t.query_device(qd, out=out[:3], flags=fl | args.flags | NN.ARB_Q_TIMING)
                st += np.array(t.timings())
            st /= max(1, args.timing_steps)
            per[vname] = {"ms_per_step": ms, "stage_ms": [round(x, 5) for x in st.tolist()]}
            log("%s %-13s: %.3f ms/step  %.3e q/s  stages %s" % (prec, vname, ms,
                                                                  args.nq / ms * 1e3, per[vname]["stage_ms"]))
        res[prec] = dict(obj=t, build_ms=build_ms, out=out, per=per)
    nc = res["mixed"]["obj"].nc
    cells = res["mixed"]["out"][3]
    valid = cells[cells < nc]
    U = int(torch.unique(valid).numel())

    # accuracy of the mixed path against the FP64 oracle on a sample
    from oracle import arbinterp_oracle as O
    rng = np.random.default_rng(7)
    idx = np.sort(rng.choice(args.nq, min(args.err_sample, args.nq), replace=False))
    o = O.tricubic(rows, "quiet", mode="both")
    ref = o.Query(q[idx].copy())
    scales = o.condition_scale(q[idx])
    err = {}
    for prec in ("mixed", "fp64"):
        c, nv, g = (x.double().cpu().numpy() for x in res[prec]["out"][:3])
        got = (c[idx], nv[idx, None], g[idx])
        e = {}
        for name, a, b, s in zip(("comps", "norm", "grad"), got, ref, scales):
            mx, rms = O.rel_err_normwise(a, b, s)
            emx, erms = O.rel_err(a, b, s)
            e[name] = {"normwise_max": mx, "normwise_rms": rms, "elementwise_max": emx,
                       "elementwise_rms": erms}
        err[prec] = e
    log("errors vs oracle:", json.dumps(err))

    # The step is timed on the library's default path ("auto").  The roofline
    # figure is for the evaluation kernel of that path: algorithmic bytes
    # N*(24 + 7*s_o) + U*4*64*s_c over its HIP-event duration.
    N = args.nq
    head = "auto" if "auto" in res["mixed"]["per"] else next(iter(res["mixed"]["per"]))
    ms = res["mixed"]["per"][head]["ms_per_step"]
    ms64 = res["fp64"]["per"][head]["ms_per_step"]
    alg_bytes = N * (24 + 7 * 4) + U * 4 * 64 * 4
    alg64 = N * (24 + 7 * 8) + U * 4 * 64 * 8
    ev_ms = res["mixed"]["per"][head]["stage_ms"][3] or ms
    ev64 = res["fp64"]["per"][head]["stage_ms"][3] or ms64
    achieved = alg_bytes / (ev_ms * 1e-3) / 1e9
    ach64 = alg64 / (ev64 * 1e-3) / 1e9

    cpu = None
    if args.cpu_chunks > 0:
        tb = time.time()
        o.allCoeffs()
        cold = time.time() - tb
        chunk = 1 << 16
        tq = time.time()
        for k in range(args.cpu_chunks):
            o.rQuery(q[k * chunk:(k + 1) * chunk].copy())
        el = time.time() - tq
        cpu = {"value": args.cpu_chunks * chunk / el, "unit": "queries/s", "cores": 1,
               "kind": "port",
               "sample": "%d x 2^16 queries of the same workload through the numpy oracle "
                         "(reference-faithful vectorised restatement), table prebuilt "
                         "(%.1f s cold fill, single thread)" % (args.cpu_chunks, cold)}
        log("cpu baseline: %.3e q/s" % cpu["value"])

    qps = N / (ms * 1e-3)
    line = {
        "metric": METRIC,
        "value": qps,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "tricubic dipole field %dx%dx%d, h=(1.0,0.8,1.25) mm, mode='both', "
                        "%d interior FP64 positions resident in HBM" % (n + (N,)),
            "precision": "mixed (FP64 cell reduction, FP32 table and evaluation)",
            "queries": N, "cells": nc, "distinct_cells_touched": U,
        },
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "evaluation kernel of the default path", "kernel_ms": ev_ms},
        "cpu_baseline": cpu,
        "fp64_plain": {"value": N / (ms64 * 1e-3), "ms_per_step": ms64,
                       "eval_kernel_ms": ev64, "roofline_achieved_GBs": ach64,
                       "frac": ach64 / HBM_PEAK_GBS},
        "speedup_vs_fp64": ms64 / ms,
        "paths": {p: res[p]["per"] for p in ("mixed", "fp64")},
        "table_build_ms": {"mixed": res["mixed"]["build_ms"], "fp64": res["fp64"]["build_ms"]},
        "rel_err_vs_fp64_oracle": err,
        "algorithmic_bytes_per_step": alg_bytes,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

"""Benchmark: value+gradient queries/s of the tricubic interpolant on one GPU.

Workload: 128^3 vector field of 6 point dipoles (h = 1.0, 0.8, 1.25 mm, N x 7
rows), mode='both' (components, |B| and grad|B|), 10^7 query positions
uniform in the interpolation interior, already resident in HBM.  One step =
one batch query of all 10^7 positions, with the outputs in the reference's
layout (comps, |B| and grad|B| arrays).  The headline path is mixed
precision (FP64 cell reduction and per-block node models, FP32 node-tile
differences and evaluation, FP32 outputs); the plain FP64 path (FP64
coefficient table, FP64 outputs) is timed beside it, and so are the other
query paths and the packed-record layouts.

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
# the bound of the metric's relative error: normwise per query and output
# kind, with the oracle's conditioning floor (oracle.rel_err_normwise)
ERR_BOUND = 1e-6
# evaluation kernel of the headline step (mixed, 3D 'both', FP32 output
# records split into the reference's arrays afterwards)
HEAD_KERNEL = "k_eval_tiles3<float, 4, 4, 4, 2>"
# bytes of one node-tile head (TileHead: 32 doubles and 4 int32)
TILE_HEAD_BYTES = 272


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
    ap.add_argument("--host-calls", type=int, default=2,
                    help="timed calls of the drop-in's numpy Query on the whole batch (PCIe "
                         "included; 0: skip)")
    ap.add_argument("--flags", type=int, default=0, help="extra ARB_Q_* flags")
    ap.add_argument("--variants", default="", help="comma list of query paths to time "
                    "(auto,direct,sorted_lds,sorted_table,sorted_fused,sorted_cell,sorted_atomic,sorted_global); default all")
    ap.add_argument("--timing-steps", type=int, default=3)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                    help="only run a plain device-to-device copy of MB MiB, 5 times (scales the "
                         "FETCH_SIZE / WRITE_SIZE counters of a profiler run), then exit")
    ap.add_argument("--workload", default="dipole3d",
                    choices=["dipole3d", "scalar3d", "moving4d", "smooth3d", "autograd4d",
                             "multifield", "point", "clustered"],
                    help="dipole3d: the headline JSON line; the others print their own line")
    ap.add_argument("--profile-only", action="store_true",
                    help="only the headline path (default flags, output arrays): --steps "
                         "queries (for counter runs; no oracle, no CPU baseline, no JSON line)")
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed",
                    help="field precision of --profile-only (fp64: the FP64 table path)")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu():
    """The host's CPU model and the cores this process may run on."""
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "cores_available": len(os.sched_getaffinity(0))}


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

    if args.copy_calibration:
        nbytes = args.copy_calibration << 20
        src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("copied %d bytes 5 times" % nbytes)
        return
    if args.profile_only:
        n = (args.grid,) * 3
        rows, q = W.dipole3d(seed=2, n=n, nq=args.nq)
        qd = torch.from_numpy(q).cuda()
        t = tricubic(rows, "quiet", mode="both", precision=args.precision)
        out = t.query_device(qd, flags=args.flags)
        for _ in range(args.steps):
            t.query_device(qd, out=out, flags=args.flags)
        torch.cuda.synchronize()
        log("profile-only: %d steps done" % args.steps)
        return

    if args.workload != "dipole3d":
        return OTHER[args.workload](args, torch)

    n = (args.grid,) * 3
    t0 = time.time()
    rows, q = W.dipole3d(seed=2, n=n, nq=args.nq)
    log("workload generated in %.1f s" % (time.time() - t0))
    qd = torch.from_numpy(q).cuda()

    from arbinterp_amd import _native as NN
    variants = {"auto": 0, "direct": NN.ARB_Q_DIRECT, "sorted_lds": NN.ARB_Q_SORTED,
                "sorted_fused": NN.ARB_Q_SORTED | NN.ARB_Q_FUSED,
                "sorted_table": NN.ARB_Q_SORTED | NN.ARB_Q_TABLE,
                "sorted_cell": NN.ARB_Q_SORTED | NN.ARB_Q_SORT_CELL,
                "sorted_atomic": NN.ARB_Q_SORTED | NN.ARB_Q_GROUP_ATOMIC,
                "sorted_global": NN.ARB_Q_SORTED_GLOBAL}
    if args.variants:
        variants = {k: v for k, v in variants.items() if k in args.variants.split(",")}
    if "auto" not in variants:
        variants = dict(auto=0, **variants)
    res = {}
    for prec in ("mixed", "fp64"):
        t = tricubic(rows, "quiet", mode="both", precision=prec)
        build_ms = one_off_build_ms(torch, t, qd)
        # output layouts: "arrays" the reference's three arrays (comps N x 3,
        # |B| N, grad N x 3; FP32 for mixed, FP64 for fp64: the drop-in's
        # default, and the headline), "p64" one 64-byte FP64 record per row,
        # "p32" one 32-byte FP32 record (mixed only)
        out = t.query_device(qd, cells=True)
        recs = {"p64": t.query_packed(qd, out_f64=True)}
        if prec == "mixed":
            recs["p32"] = t.query_packed(qd)
        torch.cuda.synchronize()
        per = {}
        for vname, fl in variants.items():
            for layout in ["arrays"] + list(recs):
                f = fl | args.flags
                if layout == "arrays":
                    fn = lambda f=f: t.query_device(qd, out=out[:3], flags=f)
                else:
                    r64 = layout == "p64"
                    fn = lambda f=f, r=recs[layout], r64=r64: t.query_packed(qd, out=r, flags=f,
                                                                            out_f64=r64)
                ms = time_steps(torch, fn, args.steps, args.warmup)
                st = np.zeros(5)
                for _ in range(args.timing_steps):
                    fn(f=f | NN.ARB_Q_TIMING)
                    st += np.array(t.timings())
                st /= max(1, args.timing_steps)
                key = vname if layout == "arrays" else vname + "_" + layout
                per[key] = {"ms_per_step": ms, "stage_ms": [round(x, 5) for x in st.tolist()]}
                log("%s %-20s: %.3f ms/step  %.3e q/s  stages %s" % (prec, key, ms,
                                                                     args.nq / ms * 1e3, per[key]["stage_ms"]))
        # the headline path's own outputs, for the accuracy check (the timing
        # loops above leave the last variant's rows in the buffers)
        out = t.query_device(qd, cells=True, out=out)
        torch.cuda.synchronize()
        host = None
        if prec == "mixed" and args.host_calls > 0:
            # the drop-in's numpy Query on the same batch: positions and
            # outputs cross PCIe (a reported rate, not the value)
            qh = q.copy()
            t.Query(qh)
            th = time.time()
            for _ in range(args.host_calls):
                t.Query(qh)
            hms = (time.time() - th) * 1e3 / args.host_calls
            host = {"ms_per_call": hms, "queries_per_s": args.nq / hms * 1e3,
                    "calls": args.host_calls}
            log("host numpy Query: %.2f ms per call" % hms)
        res[prec] = dict(obj=t, build_ms=build_ms, out=out, per=per, host=host)
    nc = res["mixed"]["obj"].nc
    cells = res["mixed"]["out"][3]
    valid = cells[cells < nc]
    U = int(torch.unique(valid).numel())
    # distinct buckets (blocks of cells) the mixed path touched, and the
    # bytes of one node tile: records of (b + 3) nodes per axis, 16 bytes each
    info = res["mixed"]["obj"].field_info()
    blk = [int(info.block[a]) for a in range(3)]
    npos = [int(info.npos[a]) for a in range(3)]
    cx = valid % npos[0]
    cy = (valid // npos[0]) % npos[1]
    cz = valid // (npos[0] * npos[1])
    nbx, nby = -(-npos[0] // blk[0]), -(-npos[1] // blk[1])
    bk = (cx // blk[0]) + nbx * ((cy // blk[1]) + nby * (cz // blk[2]))
    UB = int(torch.unique(bk).numel())
    tile_bytes = (blk[0] + 3) * (blk[1] + 3) * (blk[2] + 3) * 16 + TILE_HEAD_BYTES

    # accuracy of both paths against the FP64 oracle on a sample of rows
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
    worst = max(v["normwise_max"] for v in err["mixed"].values())
    within = bool(worst <= ERR_BOUND)

    # The step is timed on the library's default path ("auto") with the
    # reference's output layout.  roofline.achieved: the algorithmic bytes of
    # the evaluation (N*(24 + 7*s_o) + U*4*64*s_c: FP64 positions in, 7 outputs,
    # each touched cell's 4 x 64 coefficients once; FP32 s_o = s_c = 4 for the
    # mixed path) over the evaluation kernel's own HIP-event duration.
    N = args.nq
    ms = res["mixed"]["per"]["auto"]["ms_per_step"]
    ms64 = res["fp64"]["per"]["auto"]["ms_per_step"]
    st = res["mixed"]["per"]["auto"]["stage_ms"]
    st64 = res["fp64"]["per"]["auto"]["stage_ms"]
    alg_bytes = N * (24 + 7 * 4) + U * 4 * 64 * 4
    alg64 = N * (24 + 7 * 8) + U * 4 * 64 * 8
    node_bytes = N * (16 + 4 + 32) + UB * tile_bytes   # row record, row index, output record
    ev_ms = st[3] or ms
    ev64 = st64[3] or ms64
    achieved = alg_bytes / (ev_ms * 1e-3) / 1e9
    ach64 = alg64 / (ev64 * 1e-3) / 1e9
    copy_gbs = copy_bandwidth(torch, NN)
    log("copy kernel: %.1f GB/s" % copy_gbs)

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
                         "(%.1f s cold fill, single thread)" % (args.cpu_chunks, cold),
               "host": host_cpu()}
        log("cpu baseline: %.3e q/s" % cpu["value"])

    per_m = res["mixed"]["per"]
    per_d = res["fp64"]["per"]
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
            "precision": "mixed (FP64 cell reduction and tile models, FP32 node-tile "
                         "differences and evaluation)",
            "outputs": "the reference's layout: comps (N,3), |B| (N), grad|B| (N,3) as separate "
                       "FP32 arrays (query_device); the packed 64-byte FP64 and 32-byte FP32 "
                       "records are timed under paths.*_p64 / *_p32",
            "queries": N, "cells": nc, "distinct_cells_touched": U,
            "bucket_block": blk, "distinct_buckets_touched": UB,
            "node_tile_bytes": tile_bytes,
        },
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": profiled_traffic(HEAD_KERNEL),
                     "kernel": HEAD_KERNEL, "kernel_ms": ev_ms,
                     "algorithmic_bytes": alg_bytes,
                     "algorithmic_bytes_per_query": alg_bytes / N,
                     "copy_kernel_GBs": copy_gbs, "frac_of_copy": achieved / copy_gbs,
                     "node_tile_bytes": node_bytes,
                     "node_tile_achieved": node_bytes / (ev_ms * 1e-3) / 1e9,
                     "split_ms": st[4],
                     "eval_plus_split_frac": alg_bytes / ((ev_ms + st[4]) * 1e-3) / 1e9
                     / HBM_PEAK_GBS},
        "cpu_baseline": cpu,
        "accuracy": {"metric": "normwise per query and output kind: |o-r|_2 / max(|r|_2, "
                               "1e-4 |S|_2), S the oracle's condition scale; elementwise "
                               "under the same floor reported beside it",
                     "bound": ERR_BOUND, "mixed_normwise_max": worst,
                     "within_bound": within, "sample_rows": int(len(idx))},
        "fp64_plain": {"value": N / (ms64 * 1e-3), "ms_per_step": ms64,
                       "eval_kernel_ms": ev64, "split_ms": st64[4],
                       "roofline_achieved_GBs": ach64, "frac": ach64 / HBM_PEAK_GBS,
                       "algorithmic_bytes": alg64},
        "speedup_vs_fp64": ms64 / ms,
        "packed_records": {
            "mixed_p64_ms": per_m["auto_p64"]["ms_per_step"],
            "mixed_p32_ms": per_m["auto_p32"]["ms_per_step"],
            "fp64_p64_ms": per_d["auto_p64"]["ms_per_step"],
            "speedup_p64_vs_fp64_p64": per_d["auto_p64"]["ms_per_step"] / per_m["auto_p64"]["ms_per_step"],
        },
        "paths": {p: res[p]["per"] for p in ("mixed", "fp64")},
        "host_numpy_path": res["mixed"]["host"],
        "one_off_build_ms": {"mixed_node_tiles": res["mixed"]["build_ms"],
                             "fp64_table": res["fp64"]["build_ms"]},
        "rel_err_vs_fp64_oracle": err,
    }
    print(json.dumps(line), flush=True)
    if not within:
        log("FAILED: mixed normwise error %.3e exceeds the bound %.0e" % (worst, ERR_BOUND))
        sys.exit(3)


def one_off_build_ms(torch, t, qd):
    """Duration (HIP events) of the field's one-off build for its default
    query path: the node tiles of a mixed 3D field, the coefficient table
    otherwise.  Timed on a rebuild (same node values set again), so no
    allocation falls inside."""
    import ctypes
    from arbinterp_amd import _native as NN
    L = NN.lib()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    vals = torch.empty((t.field_info().ninterp, t.field_info().nnodes), dtype=torch.float64,
                       device="cuda")
    NN.check(L.arb_field_values(t._h, ctypes.c_void_p(vals.data_ptr()), st))
    tiles = t.precision == "mixed"

    def build():
        if tiles:
            t.query_device(qd[:0])
        else:
            NN.check(L.arb_build_coeffs(t._h, st))
    build()
    NN.check(L.arb_field_set_values(t._h, ctypes.c_void_p(vals.data_ptr()), st))
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    build()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1)


def copy_bandwidth(torch, NN, mib=512, reps=10):
    """GB/s of the library's plain 16-byte-vector device copy (read + write
    bytes), on the current stream."""
    import ctypes
    nb = mib << 20
    src = torch.ones(nb // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    L = NN.lib()

    def fn():
        NN.check(L.arb_device_copy(ctypes.c_void_p(dst.data_ptr()),
                                   ctypes.c_void_p(src.data_ptr()), nb, st))
    ms = time_steps(torch, fn, reps, 3)
    return 2 * nb / (ms * 1e-3) / 1e9


def profiled_traffic(kernel):
    """HBM bytes per dispatch of `kernel` from the rocprofv3 counter summaries
    under profiles/ (line "traffic_bytes_per_dispatch <kernel> = <bytes>":
    FETCH_SIZE + WRITE_SIZE of the headline workload, each scaled by the
    factor a plain copy of known size reads for that counter), or None."""
    import glob
    import re
    pat = re.compile(r"^traffic_bytes_per_dispatch (.+?) = (\d+)\s*$")
    here = os.path.dirname(os.path.abspath(__file__))
    for path in sorted(glob.glob(os.path.join(here, "profiles", "*.txt"))):
        with open(path) as fh:
            for line in fh:
                m = pat.match(line)
                if m and m.group(1) == kernel:
                    return int(m.group(2))
    return None


def _events_ms(torch, fn, steps, warmup):
    return time_steps(torch, fn, steps, warmup)


def bench_scalar3d(args, torch):
    """24x20x16 scalar field, 1e4 queries over the full box: latency of one
    batch query (device-resident and host numpy path) against the oracle."""
    from arbinterp_amd import workloads as W
    from arbinterp_amd.ARBInterp import tricubic
    from oracle import arbinterp_oracle as O
    rows, q = W.scalar3d_gaussians(seed=1)
    t = tricubic(rows, "quiet")
    t.allCoeffs()
    qd = torch.from_numpy(q).cuda()
    out = t.query_device(qd)
    dev_ms = _events_ms(torch, lambda: t.query_device(qd, out=out), args.steps, args.warmup)
    t0 = time.time()
    for _ in range(args.steps):
        t.Query(q.copy())
    host_ms = (time.time() - t0) * 1e3 / args.steps
    o = O.tricubic(rows, "quiet")
    o.allCoeffs()
    t0 = time.time()
    for _ in range(3):
        o.Query(q.copy())
    cpu_ms = (time.time() - t0) * 1e3 / 3
    print(json.dumps({"workload": "scalar3d 24x20x16, 1e4 queries (full box)",
                      "device_ms_per_query_call": dev_ms, "host_numpy_ms_per_call": host_ms,
                      "cpu_oracle_ms_per_call": cpu_ms, "queries": len(q)}))


def bench_moving4d(args, torch):
    """48^3 x 16 scalar quadcubic, 2e6 interior queries, mixed and FP64."""
    from arbinterp_amd import workloads as W
    from arbinterp_amd.ARBInterp import quadcubic
    from oracle import arbinterp_oracle as O
    rows, q = W.moving4d(seed=3)
    qd = torch.from_numpy(q).cuda()
    res = {}
    for prec in ("mixed", "fp64"):
        t = quadcubic(rows, "quiet", precision=prec)
        torch.cuda.synchronize()
        t0 = time.time()
        t.allCoeffs()
        build = (time.time() - t0) * 1e3
        out = t.query_device(qd)
        ms = _events_ms(torch, lambda: t.query_device(qd, out=out), args.steps, args.warmup)
        res[prec] = {"ms_per_step": ms, "queries_per_s": len(q) / ms * 1e3, "table_build_ms": build,
                     "out": out}
    idx = np.random.default_rng(8).choice(len(q), 20000, replace=False)
    o = O.quadcubic(rows, "quiet")
    ref = o.Query(q[idx].copy())
    sc = o.condition_scale(q[idx])
    errs = {}
    for prec in ("mixed", "fp64"):
        _, nv, g = (x.double().cpu().numpy() if x is not None else None for x in res[prec].pop("out"))
        errs[prec] = {"value": O.rel_err_normwise(nv[idx], ref[0][:, 0], sc[0][:, 0]),
                      "grad": O.rel_err_normwise(g[idx], ref[1], sc[1])}
    print(json.dumps({"workload": "quadcubic moving Gaussians 48^3x16, 2e6 queries",
                      "results": res, "rel_err_normwise_vs_oracle": errs}))


def bench_smooth3d(args, torch):
    """96^3 smooth scalar field, 2^20 FP32 positions: forward + backward of
    L = sum w U + sum |grad U|^2 through the autograd function."""
    from arbinterp_amd import workloads as W
    from arbinterp_amd.ARBInterp import tricubic
    from arbinterp_amd import autograd as AG
    rows, q, w = W.smooth3d(seed=4)
    res = {}
    for prec in ("fp64", "mixed"):
        t = tricubic(rows, "quiet", precision=prec)
        fld = AG.DifferentiableField(t, dtype=torch.float32)
        x = torch.from_numpy(q).cuda().requires_grad_(True)
        wt = torch.from_numpy(w).cuda()

        def step():
            x.grad = None
            fld.values.grad = None
            U, G = fld(x)
            (wt * U + (G * G).sum(dim=1)).sum().backward()
        ms = _events_ms(torch, step, args.steps, args.warmup)
        res[prec] = {"ms_per_forward_backward": ms, "queries": len(q)}
    print(json.dumps({"workload": "autograd 96^3 scalar, 2^20 FP32 positions", "results": res}))


def bench_autograd4d(args, torch):
    """Quadcubic moving Gaussians (48^3 x 16 nodes), 2e6 FP32 positions:
    forward alone, and forward + backward of L = sum w U + sum |grad U|^2
    through the autograd function."""
    from arbinterp_amd import workloads as W
    from arbinterp_amd.ARBInterp import quadcubic
    from arbinterp_amd import autograd as AG
    rows, q = W.moving4d(seed=3)
    w = np.random.default_rng(1).normal(size=len(q)).astype(np.float32)
    res = {}
    for prec in ("fp64", "mixed"):
        t = quadcubic(rows.copy(), "quiet", precision=prec)
        fld = AG.DifferentiableField(t, dtype=torch.float32)
        x = torch.from_numpy(q.astype(np.float32)).cuda().requires_grad_(True)
        wt = torch.from_numpy(w).cuda()

        def step(bwd=True):
            x.grad = None
            fld.values.grad = None
            U, G = fld(x)
            L = (wt * U + (G * G).sum(dim=1)).sum()
            if bwd:
                L.backward()
        fwd = _events_ms(torch, lambda: step(False), args.steps, args.warmup)
        both = _events_ms(torch, step, args.steps, args.warmup)
        res[prec] = {"ms_per_forward": fwd, "ms_per_forward_backward": both, "queries": len(q)}
    print(json.dumps({"workload": "autograd 48^3 x 16 quadcubic, 2e6 FP32 positions",
                      "results": res}))


def bench_multifield(args, torch):
    """64 uneven scalar fields in one persistent launch, then the RK4 push of
    1e6 particles for 1000 steps (dt = 1e-3) in field 0."""
    from arbinterp_amd import workloads as W
    from arbinterp_amd.ARBInterp import tricubic
    from arbinterp_amd import batch as B
    fields = W.multi_field(seed=5)
    res = {}
    for prec in ("mixed", "fp64"):
        interps = [tricubic(r, "quiet", precision=prec) for r, _ in fields]
        counts = [len(q) for _, q in fields]
        qd = torch.from_numpy(np.concatenate([q for _, q in fields])).cuda()
        bt = B.FieldBatch(interps)
        out = bt.query(qd, counts)
        ms = _events_ms(torch, lambda: bt.query(qd, counts, out=out), args.steps, args.warmup)
        # the push uses field 0 with its sign flipped: the Gaussian bumps become
        # wells, so that under a = -grad U the particles stay bound
        well = fields[0][0].copy()
        well[:, 3] *= -1.0
        t0 = tricubic(well, "quiet", precision=prec)
        lo = np.array([t0.xIntMin, t0.yIntMin, t0.zIntMin])
        hi = np.array([t0.xIntMax, t0.yIntMax, t0.zIntMax])
        rng = np.random.default_rng(11)
        x0 = lo + (hi - lo) * (0.25 + 0.5 * rng.uniform(size=(1_000_000, 3)))
        x = torch.from_numpy(x0).cuda()
        v = torch.zeros_like(x)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        B.push_rk4(t0, x, v, dt=1e-3, nsteps=1000, steps_per_launch=50, graph=True)
        e1.record()
        torch.cuda.synchronize()
        push_ms = e0.elapsed_time(e1)
        res[prec] = {"batch_ms": ms, "batch_queries": int(sum(counts)),
                     "batch_queries_per_s": sum(counts) / ms * 1e3, "push_ms_1000_steps": push_ms,
                     "push_steps_per_s": 1000 / push_ms * 1e3,
                     "push_evals_per_s": 4 * 1000 * 1e6 / push_ms * 1e3,
                     "lost_particles": int(torch.isnan(x[:, 0]).sum())}
        res[prec]["final_x"] = x.cpu().numpy()
    d = res["mixed"].pop("final_x") - res["fp64"].pop("final_x")
    m = ~np.isnan(d).any(axis=1)
    print(json.dumps({"workload": "64 uneven fields, batched; RK4 push 1e6 particles x 1000 steps",
                      "results": res, "push_mixed_vs_fp64_max_abs_dx": float(np.abs(d[m]).max())}))


def bench_point(args, torch):
    """Single-point Query on the 128^3 dipole field, mode 'both' (the per-point
    loop of examples/3D_ARBInterpExample.py:23): microseconds per call through
    the drop-in, mixed and FP64, beside the oracle's sQuery on the same points
    (which forms that cell's coefficients per call, as the reference does)."""
    from arbinterp_amd import workloads as W
    from arbinterp_amd.ARBInterp import tricubic
    from oracle import arbinterp_oracle as O
    n = (args.grid,) * 3
    rows, q = W.dipole3d(seed=2, n=n, nq=4000)
    res = {}
    for prec in ("mixed", "fp64"):
        t = tricubic(rows, "quiet", mode="both", precision=prec)
        t.allCoeffs()
        for i in range(200):
            t.Query(q[i])
        t0 = time.perf_counter()
        for i in range(len(q)):
            t.Query(q[i])
        us = (time.perf_counter() - t0) * 1e6 / len(q)
        # the same points as one-row batches (device buffers and copies per call)
        t0 = time.perf_counter()
        for i in range(500):
            t.Query(q[i:i + 1].copy())
        res[prec] = {"us_per_call": us,
                     "one_row_batch_us_per_call": (time.perf_counter() - t0) * 1e6 / 500}
        t.close()
    o = O.tricubic(rows, "quiet", mode="both")
    m = 200
    t0 = time.perf_counter()
    for i in range(m):
        o.Query(q[i])
    cpu_us = (time.perf_counter() - t0) * 1e6 / m
    print(json.dumps({"workload": "single-point Query, 128^3 dipole field, mode both",
                      "calls": len(q), "results": res,
                      "cpu_oracle_us_per_call": cpu_us, "cpu_sample_calls": m}))


def bench_clustered(args, torch):
    """128^3 dipole field, mode 'both', mixed precision, three FP32 arrays:
    the step on row sets that load a few buckets heavily (1e7 rows, made on
    the device): uniform; every row in one cell; in one bucket's span; in a
    ball over about 250 buckets; half the rows outside the box.  Each on the
    default grouping and on the LSD radix sort (ARB_Q_SORT_CELL, which also
    orders the rows by cell inside a bucket), with the stage times."""
    from arbinterp_amd import workloads as W
    from arbinterp_amd import _native as NN
    from arbinterp_amd.ARBInterp import tricubic
    rows, _ = W.dipole3d(seed=2, n=(args.grid,) * 3, nq=1)
    t = tricubic(rows, "quiet", mode="both", precision="mixed")
    f64 = dict(dtype=torch.float64, device="cuda")
    lo, hi = torch.tensor(t._lo, **f64), torch.tensor(t._hi, **f64)
    h = torch.tensor(list(t.field_info().h)[:3], **f64)
    n = args.nq
    g = torch.Generator(device="cuda")
    g.manual_seed(1)

    def u():
        return torch.rand((n, 3), generator=g, **f64)
    p0 = lo + 0.41 * (hi - lo)
    sets = {
        "uniform": lo + (hi - lo) * u(),
        "one_cell": torch.floor((p0 - lo) / h) * h + lo + 0.05 * h + 0.9 * h * u(),
        "one_bucket_span": p0 + u() * h * torch.tensor([4.0, 4.0, 2.0], **f64),
        "ball_250_buckets": p0 + u() * h * 20.0,
        "half_outside": torch.where((torch.arange(n, device="cuda") % 2 == 0)[:, None],
                                    lo + (hi - lo) * u(), lo - 0.2 * (hi - lo) * u() - 0.01),
    }
    res = {}
    out = None
    for name, q in sets.items():
        out = t.query_device(q, out=out)
        res[name] = {}
        for lab, fl in (("grouping", NN.ARB_Q_SORTED),
                        ("radix_by_cell", NN.ARB_Q_SORTED | NN.ARB_Q_SORT_CELL)):
            ms = time_steps(torch, lambda fl=fl: t.query_device(q, out=out, flags=fl),
                            args.steps, args.warmup)
            t.query_device(q, out=out, flags=fl | NN.ARB_Q_TIMING)
            res[name][lab] = {"ms_per_step": ms,
                              "stage_ms": [round(x, 5) for x in t.timings()]}
            log("%-18s %-14s %.3f ms  %s" % (name, lab, ms, res[name][lab]["stage_ms"]))
    print(json.dumps({"workload": "clustered rows, 128^3 dipole field, mode both, mixed, "
                      "%d rows" % n, "stages": "keys, grouping, -, evaluation, split",
                      "results": res}))


OTHER = {"scalar3d": bench_scalar3d, "moving4d": bench_moving4d, "smooth3d": bench_smooth3d,
         "autograd4d": bench_autograd4d,
         "multifield": bench_multifield, "point": bench_point, "clustered": bench_clustered}


if __name__ == "__main__":
    main()

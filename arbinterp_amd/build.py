"""Build libarbinterp.so (gfx950) in-tree with hipcc.

    python -m arbinterp_amd.build [-v]

Each .hip source compiles to an object under arbinterp_amd/_build/, and the
objects link into arbinterp_amd/libarbinterp.so.  Flags: no fast-math, IEEE
FP64 division, FP contraction off (every fused multiply-add is an explicit
fma() in the source), packed-FP32 instructions disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libarbinterp.so")
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["arb_capi.hip", "k_grid.hip", "k_coef.hip", "k_query.hip",
           "k_backward.hip", "k_batch.hip", "k_sort.hip", "arb_io.hip"]

CFLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-I" + os.path.join(HERE, "..", "include"),
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed (%d): %s\n%s%s" % (r.returncode, " ".join(cmd),
                                                              r.stdout, r.stderr))
    return r


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmds.append([HIPCC] + CFLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=min(8, len(cmds))) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    tmp = OUT + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-Wl,--no-undefined", "-o", tmp]
         + objs + ["-L/opt/rocm/lib", "-lamdhip64"], verbose)
    # the library must load (every symbol resolved) before it replaces the old one
    r = subprocess.run([sys.executable, "-c", "import ctypes,sys; ctypes.CDLL(sys.argv[1])", tmp],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("built library does not load:\n" + r.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

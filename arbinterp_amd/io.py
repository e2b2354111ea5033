"""Field ingestion.

`load_csv(path)` reads a numeric CSV field file the way the reference's
examples do, `np.genfromtxt(path, delimiter=',')`
(examples/3D_ARBInterpExample.py:14, examples/4D_ARBInterpExample.py:14), with
the library's host-threaded parser (`arb_csv_shape` / `arb_csv_read`): '#'
comments, blank lines skipped, empty or non-numeric fields NaN, and a
ValueError when a line's column count differs from the first line's (as
genfromtxt raises).  It needs no GPU.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _native as N


def load_csv(path) -> np.ndarray:
    """(rows, columns) float64 array of the CSV file at `path`."""
    p = os.fsencode(os.fspath(path))
    if not os.path.exists(p):
        raise OSError("%s not found." % os.fsdecode(p))
    nrows, ncols = ctypes.c_int64(), ctypes.c_int32()
    rc = N.lib().arb_csv_shape(p, ctypes.byref(nrows), ctypes.byref(ncols))
    if rc != N.ARB_OK:
        raise OSError("cannot read %s" % os.fsdecode(p))
    out = np.empty((nrows.value, ncols.value), dtype=np.float64)
    rc = N.lib().arb_csv_read(p, out.ctypes.data_as(ctypes.c_void_p), nrows.value, ncols.value)
    if rc == N.ARB_ESHAPE:
        raise ValueError("%s: every line must have %d columns" % (os.fsdecode(p), ncols.value))
    if rc != N.ARB_OK:
        raise OSError("cannot read %s" % os.fsdecode(p))
    return out

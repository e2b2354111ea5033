"""ctypes binding of libarbinterp.so (include/arbinterp.h).

The library is built in-tree (`python -m arbinterp_amd.build`).  There is no
fallback: if the library or a GPU is missing, the calls raise.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libarbinterp.so")

ARB_OK = 0
ARB_EINVAL = -1
ARB_ESHAPE = -2
ARB_ENOMEM = -3
ARB_EHIP = -4
ARB_EGRID = -5
ARB_EINDEX = -6

ARB_MODE_SCALAR = 0
ARB_MODE_VECTOR = 1
ARB_MODE_NORM = 2
ARB_MODE_BOTH = 3
ARB_MODE_VECTOR_NAN = 4
ARB_MODE_F32_NORM = 0x100

ARB_BC_NONE = 0
ARB_BC_NEUMANN = 1
ARB_BC_DIRICHLET = 2

ARB_PREC_FP64 = 0
ARB_PREC_MIXED = 1

ARB_Q_NAN_ROWS = 1
ARB_Q_CHECK = 2
ARB_Q_OUT_F64 = 4
ARB_Q_DIRECT = 8
ARB_Q_SORTED = 16
ARB_Q_SORTED_GLOBAL = 32
ARB_Q_TIMING = 64
ARB_Q_GROUP_ATOMIC = 128
ARB_Q_PACKED = 256
ARB_Q_SORT_CELL = 512
ARB_Q_FUSED = 1024
ARB_Q_TABLE = 2048
ARB_Q_REDUCE_F32 = 4096

c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_vp = ctypes.c_void_p
c_dp = ctypes.POINTER(ctypes.c_double)


class FieldInfo(ctypes.Structure):
    _fields_ = [
        ("ndim", c_i32), ("mode", c_i32), ("precision", c_i32), ("ninterp", c_i32),
        ("has_comps", c_i32), ("has_norm", c_i32), ("table_built", c_i32), ("coord_f32", c_i32),
        ("n", c_i64 * 4), ("npos", c_i64 * 4), ("h", ctypes.c_double * 4),
        ("int_min", ctypes.c_double * 4), ("int_max", ctypes.c_double * 4),
        ("nc", c_i64), ("nnodes", c_i64), ("table_bytes", c_i64),
        ("packed_width", c_i32), ("boundary", c_i32), ("block", c_i32 * 4),
        ("tiles_bytes", c_i64), ("on_the_fly", c_i32), ("reserved", c_i32),
    ]


# (name, restype, argtypes) for every symbol declared in include/arbinterp.h
SIGNATURES = [
    ("arb_field_create", c_i32, [c_vp, c_i64, c_i32, c_i64, c_i32, c_i32, c_i32, c_vp,
                                 ctypes.POINTER(c_vp)]),
    ("arb_field_info_get", c_i32, [c_vp, ctypes.POINTER(FieldInfo)]),
    ("arb_field_set_boundary", c_i32, [c_vp, c_i32, c_vp]),
    ("arb_build_coeffs", c_i32, [c_vp, c_vp]),
    ("arb_set_table_limit", c_i32, [c_vp, c_i64]),
    ("arb_query", c_i32, [c_vp, c_vp, c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_i32, c_vp]),
    ("arb_query_noalloc", c_i32, [c_vp, c_vp, c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_i32, c_vp]),
    ("arb_query_point", c_i32, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    ("arb_table_copy", c_i32, [c_vp, c_i64, c_i64, c_vp, c_vp]),
    ("arb_table_load", c_i32, [c_vp, c_i64, c_i64, c_vp, c_vp]),
    ("arb_field_values", c_i32, [c_vp, c_vp, c_vp]),
    ("arb_field_set_values", c_i32, [c_vp, c_vp, c_vp]),
    ("arb_query_backward", c_i32, [c_vp, c_vp, c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp]),
    ("arb_batch_create", c_i32, [ctypes.POINTER(c_vp), c_i32, c_vp, ctypes.POINTER(c_vp)]),
    ("arb_batch_query", c_i32, [c_vp, c_vp, ctypes.POINTER(c_i64), c_i64, c_vp, c_vp, c_vp,
                                c_i32, c_vp]),
    ("arb_batch_set_tuning", c_i32, [c_vp, c_i32, c_i32, c_i32]),
    ("arb_batch_destroy", None, [c_vp]),
    ("arb_push_rk4", c_i32, [c_vp, c_vp, c_vp, c_i64, ctypes.c_double, c_i64, c_i32, c_i32, c_vp]),
    ("arb_last_timings", c_i32, [c_vp, ctypes.POINTER(ctypes.c_float)]),
    ("arb_device_copy", c_i32, [c_vp, c_vp, c_i64, c_vp]),
    ("arb_set_tuning", c_i32, [c_vp, ctypes.POINTER(c_i32), c_i32]),
    ("arb_operator_matrix", c_i32, [c_i32, c_i32, c_vp]),
    ("arb_csv_shape", c_i32, [ctypes.c_char_p, ctypes.POINTER(c_i64), ctypes.POINTER(c_i32)]),
    ("arb_csv_read", c_i32, [ctypes.c_char_p, c_vp, c_i64, c_i32]),
    ("arb_field_destroy", None, [c_vp]),
    ("arb_last_error", ctypes.c_char_p, []),
    ("arb_version", ctypes.c_char_p, []),
]

_lib = None


def lib():
    """Load libarbinterp.so once; raise with a build hint if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libarbinterp.so is not built; run `python -m arbinterp_amd.build` "
                           "(hipcc, gfx950)")
    L = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def last_error() -> str:
    return lib().arb_last_error().decode(errors="replace")


def check(rc: int, what: str = "") -> None:
    """Map an ARB_E* status to the Python exception the reference raises."""
    if rc == ARB_OK:
        return
    msg = last_error()
    if what:
        msg = "%s: %s" % (what, msg)
    if rc == ARB_ESHAPE:
        raise SystemExit(last_error())
    if rc == ARB_EINDEX:
        raise IndexError(msg)
    if rc == ARB_ENOMEM:
        raise MemoryError(msg)
    if rc in (ARB_EINVAL, ARB_EGRID):
        raise ValueError(msg)
    raise RuntimeError(msg)

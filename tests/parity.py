"""Shared parity helpers: run a Query like the reference's callers do and
compare outputs, NaN patterns, mutated query arrays and exceptions with a
golden fixture.

Error metric (oracle.rel_err): |got - ref| / max(|ref|, 1e-4 * S) with S the
oracle's condition scale sum_i |beta_i| sum_j |A_ij f_j|.
"""
import numpy as np

from oracle import arbinterp_oracle as O


def run_query(obj, q):
    """obj.Query on a copy of q; returns (("ok", outputs) | ("exc", name), q_after)."""
    qq = np.array(q, copy=True)
    try:
        out = obj.Query(qq)
    except Exception as e:  # noqa: BLE001 - the exception type is compared
        return ("exc", type(e).__name__), qq
    if not isinstance(out, tuple):
        out = (out,)
    return ("ok", tuple(np.asarray(o, dtype=np.float64) for o in out)), qq


def same_nan_array(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.shape != b.shape:
        return False
    na, nb = np.isnan(a), np.isnan(b)
    return np.array_equal(na, nb) and np.array_equal(a[~na], b[~nb])


def compare_outputs(res, qafter, g, prefix, scale_obj, query, tol, label=""):
    """Assert that `res` (from run_query) matches the fixture entries under
    `prefix`; `scale_obj` is an oracle object giving condition scales."""
    kind, val = res
    tag = label or prefix
    if prefix + "exc" in g.files:
        assert kind == "exc", "%s: expected %s, got outputs" % (tag, g[prefix + "exc"])
        assert val == str(g[prefix + "exc"]), "%s: %s vs %s" % (tag, val, g[prefix + "exc"])
    else:
        assert kind == "ok", "%s: raised %s" % (tag, val)
        n_ref = len([k for k in g.files if k.startswith(prefix + "out")])
        assert len(val) == n_ref, tag
        q2 = np.atleast_2d(np.asarray(query, dtype=np.float64))
        scales = scale_obj.condition_scale(q2)
        if not isinstance(scales, tuple):
            scales = (scales,)
        for i in range(n_ref):
            ref = g[prefix + "out%d" % i]
            got = val[i]
            assert got.shape == ref.shape, "%s out%d shape %s vs %s" % (tag, i, got.shape, ref.shape)
            assert np.array_equal(np.isnan(got), np.isnan(ref)), "%s out%d NaN pattern" % (tag, i)
            if np.isnan(ref).all():
                continue
            sc = np.asarray(scales[i], dtype=np.float64).reshape(ref.shape)
            mx, _ = O.rel_err(got, ref, sc)
            assert mx <= tol, "%s out%d rel err %.3e > %.1e" % (tag, i, mx, tol)
    assert same_nan_array(qafter, g[prefix + "qafter"]), "%s: mutated query differs" % tag

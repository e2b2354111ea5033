"""The C-ABI library: loads without a GPU, exports every declared symbol, and
its host-only entry points agree with the oracle."""
import ctypes
import os
import re

import numpy as np
import pytest

from arbinterp_amd import _native as N
from oracle import arbinterp_oracle as O

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "include", "arbinterp.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(arb_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = N.lib()
    names = declared_functions()
    assert len(names) >= 8
    for name in names:
        assert hasattr(L, name), name
    bound = {s[0] for s in N.SIGNATURES}
    assert set(names) <= bound, set(names) - bound


def test_version_string():
    assert b"gfx950" in N.lib().arb_version()


@pytest.mark.parametrize("d", [3, 4])
def test_operator_matrices_match_oracle(d):
    K = 4 ** d
    refs = [O.make_B(d), np.round(np.linalg.inv(O.make_B(d))), O.make_D(d), O.make_A(d)]
    for which, ref in enumerate(refs):
        out = np.zeros((K, K))
        assert N.lib().arb_operator_matrix(d, which, out.ctypes.data_as(ctypes.c_void_p)) == 0
        assert np.array_equal(out, ref), which


def test_argument_errors_without_gpu():
    L = N.lib()
    assert L.arb_operator_matrix(5, 0, None) == N.ARB_EINVAL
    h = ctypes.c_void_p()
    assert L.arb_field_create(None, 10, 4, 4, 3, 0, 0, None, None) == N.ARB_EINVAL
    assert L.arb_field_create(None, 10, 5, 5, 3, 0, 0, None, ctypes.byref(h)) == N.ARB_ESHAPE
    assert "Input not shaped" in N.last_error()
    with pytest.raises(SystemExit):
        N.check(N.ARB_ESHAPE)
    with pytest.raises(IndexError):
        N.check(N.ARB_EINDEX)


def test_packed_record_layout_documented():
    """Record widths: [comps(3)] [value, grad(d)] padded to 4 or 8 elements."""
    txt = open(HEADER).read()
    assert "ARB_Q_PACKED" in txt and "packed_width" in txt
    import ctypes as C
    assert any(f[0] == "packed_width" for f in N.FieldInfo._fields_)
    # 3D both: 3 + 1 + 3 = 7 -> 8; 3D scalar: 4 -> 4; 4D both: 8; 4D scalar: 5 -> 8; vector: 4
    width = lambda comps, norm, d: 4 if (3 * comps + (1 + d) * norm) <= 4 else 8
    assert (width(1, 1, 3), width(0, 1, 3), width(1, 1, 4), width(0, 1, 4), width(1, 0, 3)) == \
        (8, 4, 8, 8, 4)
    assert C.sizeof(N.FieldInfo) % 8 == 0


def test_reference_import_paths_bind_gpu_classes():
    """`from ARBInterp import ARBInterp` (the reference's tests/test_ARBInterp.py:1)
    and `from ARBTools.ARBInterp import tricubic, quadcubic` (its examples)
    give the GPU classes (importing needs no GPU; constructing one does)."""
    from ARBInterp import ARBInterp
    from ARBTools.ARBInterp import quadcubic, tricubic
    import arbinterp_amd.ARBInterp as G
    assert ARBInterp.tricubic is G.tricubic is tricubic
    assert ARBInterp.quadcubic is G.quadcubic is quadcubic
    assert ARBInterp.__version__ == "1.8"

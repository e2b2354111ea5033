"""Field ingestion: arbinterp_amd.io.load_csv against np.genfromtxt(path,
delimiter=','), the call the reference's examples use to load fields
(examples/3D_ARBInterpExample.py:14).  Host only."""
import numpy as np
import pytest

from arbinterp_amd import workloads as W
from arbinterp_amd.io import load_csv


def _same(a, b):
    return a.shape == b.shape and np.array_equal(a, b, equal_nan=True) and \
        np.array_equal(np.signbit(a), np.signbit(b))


def test_field_file_roundtrip(tmp_path):
    """A shuffled 3D vector field written as the examples' files are (one node
    per line, x, y, z, Bx, By, Bz), in several number formats."""
    rows, _ = W.dipole3d(seed=2, n=(12, 11, 10), nq=1)
    rng = np.random.default_rng(0)
    for k, fmt in enumerate(("%.17g", "%.6e", "%.10f", "% .8g")):
        p = tmp_path / ("field%d.csv" % k)
        np.savetxt(p, rows[rng.permutation(len(rows))], delimiter=",", fmt=fmt)
        assert _same(load_csv(p), np.genfromtxt(p, delimiter=","))


def test_grammar_edge_cases(tmp_path):
    p = tmp_path / "edge.csv"
    p.write_text("# header comment\n1.5, 2e3 ,-0.0\n\n 4,nan, inf # tail\r\n"
                 "-inf,+7.25e-3,abc\n8,,1e400\n9, 0x1p3, 1-2\n.5,5.,-1E-3")
    assert _same(load_csv(p), np.genfromtxt(p, delimiter=","))


def test_ragged_rows_raise(tmp_path):
    p = tmp_path / "ragged.csv"
    p.write_text("1,2,3\n4,5\n")
    with pytest.raises(ValueError):
        np.genfromtxt(p, delimiter=",")
    with pytest.raises(ValueError):
        load_csv(p)


def test_missing_file(tmp_path):
    with pytest.raises(OSError):
        load_csv(tmp_path / "nope.csv")


def test_single_row_and_column(tmp_path):
    p = tmp_path / "one.csv"
    p.write_text("3.25\n")
    got = load_csv(p)
    assert got.shape == (1, 1) and got[0, 0] == 3.25

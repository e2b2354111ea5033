"""The reference's import paths and its examples' call sequence.

Callers of the reference import `from ARBInterp import ARBInterp`
(tests/test_ARBInterp.py:1) or `from ARBTools.ARBInterp import tricubic /
quadcubic` (examples/3D_ARBInterpExample.py:2, 4D_ARBInterpExample.py:2).
Both resolve to the GPU classes here.  The examples load a field with
np.genfromtxt(path, delimiter=','), build the interpolator (scalar field
without a mode, then mode='both'), query coords[3] as a single point and
the 20-row coords array as a batch (3D example :14-50, 4D example :14-51).
That sequence is replayed on synthetic CSV fields (the reference's example
fields are not in its repository) and compared with the oracle.
"""
import numpy as np
import pytest

from arbinterp_amd import workloads as W
from oracle import arbinterp_oracle as O

pytestmark = pytest.mark.gpu
TOL64 = 1e-12


def _csv_field(tmp_path, name, d, ncomp, seed):
    """A field over [-3e-3, 3e-3]^3 (and t over [-5e-6, 5e-6]) written as the
    examples' CSV: one row per node, coordinates then values."""
    rng = np.random.default_rng(seed)
    n = [11, 10, 12, 8][:d]
    ext = [3e-3, 3e-3, 3e-3, 5e-6][:d]
    axes = [np.linspace(-ext[a], ext[a], n[a]) for a in range(d)]
    mesh = np.meshgrid(*axes, indexing="ij")
    vals = []
    for c in range(ncomp):
        ctr = rng.uniform(-1, 1, size=d) * np.array(ext) * 0.5
        r2 = sum(((mesh[a] - ctr[a]) / ext[a]) ** 2 for a in range(d))
        vals.append((1.0 + c) * np.exp(-r2) + 0.1 * c)
    rows = W.grid_rows(axes, vals, rng)
    path = tmp_path / name
    np.savetxt(path, rows, delimiter=",", fmt="%.17g")
    return str(path)


def _same(a, b, scale, tag):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    assert a.shape == b.shape, tag
    assert np.array_equal(np.isnan(a), np.isnan(b)), tag
    mx, _ = O.rel_err(a.reshape(b.shape), b, np.asarray(scale).reshape(b.shape))
    assert mx <= TOL64, (tag, mx)


def _example_sequence(cls_gpu, cls_ref, d, tmp_path, capsys):
    coords = np.zeros((20, d))
    for a in range(3):
        coords[:, a] = np.linspace(-2e-3, 2e-3, 20)
    if d == 4:
        coords[:, 3] = np.linspace(-3e-6, 3e-6, 20)
    for ncomp, kw in ((1, {}), (3, {"mode": "both"})):
        path = _csv_field(tmp_path, "f%d_%d.csv" % (d, ncomp), d, ncomp, seed=d + ncomp)
        field = np.genfromtxt(path, delimiter=",")
        run = cls_gpu(field, **kw)
        banner_gpu = capsys.readouterr().out
        ref = cls_ref(field.copy(), **kw)
        banner_ref = capsys.readouterr().out
        assert banner_gpu == banner_ref and banner_gpu.startswith("---")
        one = run.Query((coords[3]))
        one_ref = ref.Query((coords[3].copy()))
        scales = ref.condition_scale(coords[3:4].copy())
        for k, (a, b, s) in enumerate(zip(one, one_ref, scales)):
            _same(np.atleast_1d(a), np.atleast_1d(b), s, "single %d" % k)
        comps = run.Query(coords)
        comps_ref = ref.Query(coords.copy())
        scales = ref.condition_scale(coords.copy())
        for k, (a, b, s) in enumerate(zip(comps, comps_ref, scales)):
            _same(a, b, s, "batch %d" % k)
        run.close()


def test_3d_example_sequence(tmp_path, capsys):
    from ARBTools.ARBInterp import tricubic
    _example_sequence(tricubic, O.tricubic, 3, tmp_path, capsys)


def test_4d_example_sequence(tmp_path, capsys):
    from ARBTools.ARBInterp import quadcubic
    _example_sequence(quadcubic, O.quadcubic, 4, tmp_path, capsys)

"""Host-side checks of the C ABI and the setup helpers (CPU only, no GPU compute).

- libsssi.so loads and exports every entry point declared in include/sssi.h,
  and the ctypes table covers each of them;
- argument validation rejects bad inputs before any device work;
- setup helpers follow the reference's MATLAB helpers (extBoundary, ricker,
  dCoef, layerModelGenerator).
"""
import os
import re

import numpy as np
import pytest

import oracle
import sssi_amd
from sssi_amd import _lib, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    text = open(os.path.join(ROOT, "include", "sssi.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sssi_[A-Za-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    names = declared_entry_points()
    assert len(names) > 20
    for n in names:
        assert hasattr(lib, n), f"{n} not exported"
        assert n in _lib.SIGNATURES, f"{n} missing from the ctypes table"


def test_version_and_error_text():
    lib = _lib.lib()
    assert lib.sssi_version() >= 100
    assert isinstance(lib.sssi_last_error(), bytes)


@pytest.mark.parametrize("kwargs,frag", [
    (dict(N=9), "nDiffOrder"),
    (dict(N=0), "nDiffOrder"),
    (dict(nB=16), "2*nBoundary"),
    (dict(nB=-1), "nBoundary"),
    (dict(dt=0.0), "positive"),
    (dict(nt=0), "nt"),
])
def test_argument_validation_happens_before_device_work(kwargs, frag):
    args = dict(N=2, nB=4, dt=1e-3, nt=10)
    args.update(kwargs)
    v = np.full((20, 30), 2000.0)
    with pytest.raises(sssi_amd.SssiError) as e:
        sssi_amd.fwd_cpml_2d_points(v, args["nt"], args["N"], args["nB"], 10.0, 10.0, args["dt"],
                                    [0], [5], np.ones(max(args["nt"], 1)))
    assert e.value.code == -1 and frag in str(e.value)


def test_mex_signature_shape_checks():
    v = np.full((10, 12), 2000.0)
    with pytest.raises(ValueError):
        sssi_amd.fwdTimeCpmlFor2dAw(v, np.zeros((10, 11, 5)), 2, 2, 10.0, 10.0, 1e-3)
    with pytest.raises(ValueError):
        sssi_amd.rvsTimeCpmlFor2dAw(v, np.zeros((11, 5)), 2, 2, 10.0, 10.0, 1e-3)


def test_extBoundary_2d_and_3d():
    V = np.arange(12.0).reshape(3, 4)
    E = synth.extBoundary(V, 2, 2)
    assert E.shape == (5, 8)
    assert np.array_equal(E[:3, 2:6], V)
    assert np.all(E[:3, 0] == V[:, 0]) and np.all(E[:3, -1] == V[:, -1])
    assert np.array_equal(E[3], E[2]) and np.array_equal(E[4], E[2])   # bottom replicates
    V3 = np.random.default_rng(0).random((3, 4, 5))
    E3 = synth.extBoundary(V3, 2, 3)
    assert E3.shape == (5, 8, 9)
    assert np.array_equal(E3[:3, 2:6, 2:7], V3)
    assert np.array_equal(E3[:3, 2:6, 0], V3[:, :, 0])


def test_ricker_matches_formula():
    f, n, dt = 20.0, 200, 1e-3
    w = synth.ricker(f, n, dt)
    t = np.arange(n) * dt
    tau = t - 1 / f
    ref = (1 - 2 * np.pi ** 2 * f ** 2 * tau ** 2) * np.exp(-np.pi ** 2 * f ** 2 * tau ** 2)
    np.testing.assert_allclose(w, ref, rtol=1e-13, atol=1e-15)
    assert w[50] == pytest.approx(1.0)   # peak at t0 = 1/f


def test_dcoef_host_matches_oracle():
    for N in range(1, 9):
        assert np.array_equal(synth.dCoef(N), oracle.dcoef(N))


def test_layered_models_are_seeded():
    a = synth.layered_2d(50, 60, 5, 1500, 4500, seed=3)
    b = synth.layered_2d(50, 60, 5, 1500, 4500, seed=3)
    assert np.array_equal(a, b) and len(np.unique(a)) == 5
    assert a.min() >= 1500 and a.max() <= 4500
    c = synth.cake_3d_dipping(20, 30, 30, 8, 1500, 4500, seed=1)
    assert c.shape == (20, 30, 30) and c.min() >= 1500 and c.max() <= 4500
    L = synth.layerModelGenerator([1500, 2000], [3, 2], 4)
    assert L.shape == (5, 4) and np.all(L[:3] == 1500) and np.all(L[3:] == 2000)


def test_cfl_number():
    # config with v dt/dx = 0.45 and N = 2 in 2-D: 0.45 * sqrt(2) * (9/8 + 1/24)
    c = synth.cfl_number(4500.0, 1e-3, [10.0, 10.0], 2)
    assert c == pytest.approx(0.45 * np.sqrt(2) * (9 / 8 + 1 / 24))

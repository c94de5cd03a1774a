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


def _rc(name, *args):
    lib = _lib.lib()
    rc = getattr(lib, name)(*args)
    return rc, lib.sssi_last_error().decode()


@pytest.mark.parametrize("bad", [np.nan, np.inf, 0.0, -1500.0])
def test_velocity_validated_before_device_work(bad):
    """A NaN/Inf or non-positive velocity is rejected (SSSI_EINVAL) by every
    entry point that takes a model, before any device call (none runs here)."""
    from sssi_amd._lib import _dp, _f64
    import ctypes
    nz, nx, nt = 20, 30, 8
    v = np.full((nz, nx), 2000.0)
    v[7, 11] = bad
    vf = _f64(v)
    w = np.ones(nt)
    iz = np.zeros(1, np.int32)
    ix = np.full(1, 5, np.int32)
    I = ctypes.POINTER(ctypes.c_int)
    data = np.zeros((nx, nt), order="F")
    calls = [
        ("sssi_fwd_cpml_2d_points", _dp(vf), nz, nx, nt, 1, iz.ctypes.data_as(I),
         ix.ctypes.data_as(I), _dp(w), 2, 4, 10.0, 10.0, 1e-3, _dp(data), None, 1),
        ("sssi_fwd_sponge_2d_points", _dp(vf), nz, nx, nt, 1, iz.ctypes.data_as(I),
         ix.ctypes.data_as(I), _dp(w), 4, 10.0, 10.0, 1e-3, _dp(data), None, 1),
        ("sssi_rvs_cpml_2d_ex", _dp(vf), nz, nx, _dp(_f64(data)), nt, 2, 4, 10.0, 10.0, 1e-3,
         None, None, 1),
        ("sssi_prop2d_create", _dp(vf), nz, nx, 0, 2, 4, 10.0, 10.0, 1e-3,
         ctypes.byref(ctypes.c_void_p())),
        ("sssi_rtm2d_create", _dp(vf), _dp(vf), nz, nx, 2, 4, 10.0, 10.0, 1e-3, nt, 1, 1, 1e-3,
         ctypes.byref(ctypes.c_void_p())),
    ]
    for name, *args in calls:
        rc, msg = _rc(name, *args)
        assert rc == -1, (name, rc, msg)
        assert "finite and positive" in msg and "element 227" in msg  # column-major: 11 * 20 + 7, (name, msg)
    v3 = np.full((6, 8, 10), 2000.0)
    v3[1, 2, 3] = bad
    rc, msg = _rc("sssi_prop3d_create", _dp(_f64(v3)), 6, 8, 10, 2, 2, 10.0, 10.0, 10.0, 1e-3,
                  ctypes.byref(ctypes.c_void_p()))
    assert rc == -1 and "finite and positive" in msg


def test_3d_rejects_boundary_wider_than_a_single_y_slice():
    """A 3-D call with ny = 1 (a 2-D model) and nBoundary >= 1 has no room for
    the two y strips; the reference errors on it, and so does every 3-D entry."""
    import ctypes
    from sssi_amd._lib import _dp, _f64
    v = _f64(np.full((12, 16, 1), 2000.0))
    rc, msg = _rc("sssi_prop3d_create", _dp(v), 12, 16, 1, 2, 2, 10.0, 10.0, 10.0, 1e-3,
                  ctypes.byref(ctypes.c_void_p()))
    assert rc == -1 and "ny" in msg
    with pytest.raises(sssi_amd.SssiError) as e:
        sssi_amd.fwd_cpml_3d_points(np.full((12, 16, 1), 2000.0), 4, 2, 1, 10.0, 10.0, 10.0, 1e-3,
                                    [0], [8], [0], np.ones(4))
    assert e.value.code == -1 and "ny" in str(e.value)


def test_sponge_grid_capped_below_2gib():
    """The sponge kernel drops out-of-model stores with a 2^31 byte offset, which
    is out of range only below 2 GiB: larger padded grids are refused before
    anything is read or allocated."""
    import ctypes
    from sssi_amd._lib import _dp
    v = np.full(4, 2000.0)  # never read: the size check comes first
    nz, nx = 16384, 32704
    rc, msg = _rc("sssi_prop2d_create", _dp(v), nz, nx, 1, 1, 4, 10.0, 10.0, 1e-3,
                  ctypes.byref(ctypes.c_void_p()))
    assert rc == -1 and "2 GiB" in msg
    rc, msg = _rc("sssi_fwd_sponge_2d_points", _dp(v), nz, nx, 4, 0, None, None, None, 4, 10.0,
                  10.0, 1e-3, None, None, 1)
    assert rc == -1 and "2 GiB" in msg


def test_unstable_status_code_is_exported():
    text = open(os.path.join(ROOT, "include", "sssi.h")).read()
    assert re.search(r"#define SSSI_EUNSTABLE \(-6\)", text)
    src = open(os.path.join(ROOT, "sssi_amd", "csrc", "prop2d.hip")).read()
    assert "SSSI_EUNSTABLE" in src  # raised by the watchdog check

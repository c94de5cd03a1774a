"""sssi_fwd_cpml_2d_points_noalloc, the entry point that allocates nothing,
against sssi_fwd_cpml_2d_points (the allocating drop-in it mirrors), through
the C ABI: seismograms and strided snapshots equal bit for bit on two calls in
a row, the second on a smaller grid with fewer steps (the device array keeps
the first call's contents, which the second must not see); a problem above the
array's size is refused with SSSI_EINVAL and a message naming the limit."""
import ctypes

import numpy as np
import pytest

from sssi_amd import _lib, synth
from sssi_amd._lib import _dp, _f64

pytestmark = pytest.mark.gpu

_I = ctypes.POINTER(ctypes.c_int)


def _call(name, v, nt, N, nB, d, dt, iz, ix, w, stride):
    v = _f64(v)
    nz, nx = v.shape
    iz = np.ascontiguousarray(iz, dtype=np.int32)
    ix = np.ascontiguousarray(ix, dtype=np.int32)
    w = np.ascontiguousarray(w, dtype=np.float64).reshape(nt, iz.size)
    data = np.zeros((nx, nt), order="F")
    snap = np.zeros((nz, nx, (nt + stride - 1) // stride), order="F")
    lib = _lib.lib()
    rc = getattr(lib, name)(_dp(v), nz, nx, nt, iz.size, iz.ctypes.data_as(_I),
                            ix.ctypes.data_as(_I), _dp(w), N, nB, d, d, dt, _dp(data), _dp(snap),
                            stride)
    assert rc == 0, (name, rc, lib.sssi_last_error().decode())
    return data, snap


@pytest.mark.parametrize("second", [(90, 170, 300, 3, 9, 7), (40, 64, 120, 5, 4, 1)])
def test_noalloc_equals_allocating_entry_twice(second):
    print(f"noalloc: first 120x240 nt 1000, then {second}", flush=True)
    cases = [(120, 240, 1000, 2, 20, 10), second]
    for q, (nz, nx, nt, N, nB, stride) in enumerate(cases):
        v = synth.layered_2d(nz, nx, 5, 1500.0, 4000.0, seed=31 + q)
        dt = 0.4 * 10.0 / 4000.0 / np.sqrt(2) / float(np.sum(np.abs(synth.dCoef(N))))
        w = synth.ricker(20.0, nt, dt)
        iz, ix = [0, nz // 3, nz // 3], [nx // 2, 5, 5]   # coincident points add
        ww = np.stack([w, 0.5 * w, -0.25 * w], axis=1)
        args = (v, nt, N, nB, 10.0, dt, iz, ix, ww, stride)
        d_ref, s_ref = _call("sssi_fwd_cpml_2d_points", *args)
        d_na, s_na = _call("sssi_fwd_cpml_2d_points_noalloc", *args)
        assert np.any(d_ref != 0)
        assert np.array_equal(d_na, d_ref), q
        assert np.array_equal(s_na, s_ref), q


def test_noalloc_refuses_problems_above_its_array():
    v = np.full((2000, 3000), 2000.0)   # three padded levels alone need 77 MB
    lib = _lib.lib()
    data = np.zeros((3000, 10), order="F")
    iz = np.zeros(1, np.int32)
    ix = np.zeros(1, np.int32)
    w = np.ones(10)
    rc = lib.sssi_fwd_cpml_2d_points_noalloc(_dp(_f64(v)), 2000, 3000, 10, 1,
                                             iz.ctypes.data_as(_I), ix.ctypes.data_as(_I), _dp(w),
                                             4, 20, 10.0, 10.0, 1e-3, _dp(data), None, 1)
    assert rc == -1
    assert "64 MiB" in lib.sssi_last_error().decode()

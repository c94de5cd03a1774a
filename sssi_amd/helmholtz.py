"""Frequency-domain CPML Helmholtz over the C ABI: freqCpmlFor2dAw
(src/freqCpmlFor2dAw.m) with the MATLAB function's arguments and outputs.

  A, snapshot = freqCpmlFor2dAw(model, source, w, nDiffOrder, nBoundary, dz, dx)

model is the squared slowness (nz, nx); source is complex, (nz*nx, nsrc) or
(nz, nx[, nsrc]) (cell (i, j) is row j*nz + i, MATLAB's reshape); A is the
impedance matrix as a scipy.sparse CSR matrix (MATLAB returns a sparse
matrix); snapshot = A \\ (-source) has the shape of source.  nargout=1 returns
A alone; an empty source gives snapshot None (MATLAB's [])."""
import ctypes

import numpy as np

from ._lib import _call, _dp, _f64


def _cplx_ptr(a):
    return a.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def impedance_matrix(model, w, nDiffOrder, nBoundary, dz, dx):
    import scipy.sparse as sp
    m = _f64(model)
    nz, nx = m.shape
    nnz = ctypes.c_int64()
    _call("sssi_freq_cpml_2d_matrix", _dp(m), nz, nx, float(w), int(nDiffOrder), int(nBoundary),
          float(dz), float(dx), ctypes.byref(nnz), None, None, None)
    rp = np.zeros(nz * nx + 1, dtype=np.int64)
    col = np.zeros(nnz.value, dtype=np.int64)
    val = np.zeros(nnz.value, dtype=np.complex128)
    P = ctypes.POINTER(ctypes.c_int64)
    _call("sssi_freq_cpml_2d_matrix", _dp(m), nz, nx, float(w), int(nDiffOrder), int(nBoundary),
          float(dz), float(dx), ctypes.byref(nnz), rp.ctypes.data_as(P), col.ctypes.data_as(P),
          _cplx_ptr(val))
    return sp.csr_matrix((val, col, rp), shape=(nz * nx, nz * nx))


def solve(model, source, w, nDiffOrder, nBoundary, dz, dx):
    """snapshot = A \\ (-source) on the device; shape of source."""
    m = _f64(model)
    nz, nx = m.shape
    src = np.asarray(source, dtype=np.complex128)
    S = src.reshape(nz * nx, -1, order="F")
    ns = S.shape[1]
    s1 = np.ascontiguousarray(S.ravel(order="F"))
    u1 = np.zeros(nz * nx * ns, dtype=np.complex128)
    _call("sssi_freqCpmlFor2dAw", _dp(m), nz, nx, _cplx_ptr(s1), ns, float(w), int(nDiffOrder),
          int(nBoundary), float(dz), float(dx), _cplx_ptr(u1))
    return u1.reshape(src.shape, order="F")


def freqCpmlFor2dAw(model, source, w, nDiffOrder, nBoundary, dz, dx, nargout=2):
    A = impedance_matrix(model, w, nDiffOrder, nBoundary, dz, dx)
    if nargout < 2:
        return A
    if source is None or np.size(source) == 0:
        return A, None
    return A, solve(model, source, w, nDiffOrder, nBoundary, dz, dx)


def last_timing():
    """(factorisation ms, solve ms) of this thread's last solve (HIP events)."""
    f, s = ctypes.c_double(), ctypes.c_double()
    _call("sssi_freq_last_timing", ctypes.byref(f), ctypes.byref(s))
    return f.value, s.value

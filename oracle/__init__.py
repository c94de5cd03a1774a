"""FP64 CPU oracle for the SSSI acoustic propagators -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The product
path (``sssi_amd`` over ``libsssi.so``) never imports it.

Arrays use the reference's MATLAB index order (numpy arrays indexed like the
MATLAB ones): ``v[iz, ix]``, ``source[iz, ix, it]``, ``data[ix, it]``,
``snapshot[iz, ix, it]``; 3-D ``v[iz, ix, iy]``, ``data[ix, iy, it]``.
Internally everything is passed column-major (Fortran order) as the reference
MEX functions receive it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libsssi_oracle.so")
_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_i64 = ctypes.c_int64


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_dcoef.argtypes = [ctypes.c_int, _D]
        L.orc_cpml_profiles_2d.argtypes = [_D, _i64, _i64, ctypes.c_int, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double, _D, _D]
        L.orc_cpml_profiles_3d.argtypes = [_D, _i64, _i64, _i64, ctypes.c_int, ctypes.c_double,
                                           _D, _D, _D]
        L.orc_fwd_cpml_2d.argtypes = [_D, _i64, _i64, _i64, _D, ctypes.c_int, _I, _I, _D,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double, _D, _D, _i64]
        L.orc_rvs_cpml_2d.argtypes = [_D, _i64, _i64, _i64, _D, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      _D, _D, _i64]
        L.orc_fwd_cpml_3d.argtypes = [_D, _i64, _i64, _i64, _i64, _D, ctypes.c_int, _I, _I, _I,
                                      _D, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      _D, _D, _i64]
        L.orc_fwd_sponge_2d.argtypes = [_D, _i64, _i64, _i64, _D, ctypes.c_int, _I, _I, _D,
                                        ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, _D, _D, _i64]
        L.orc_rvs_sponge_2d.argtypes = [_D, _i64, _i64, _i64, _D, ctypes.c_int,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        _D, _D, _i64]
        L.orc_rtm_shot_cpml_2d.argtypes = [_D, _D, _i64, _i64, _i64, ctypes.c_int, _I, _I, _D,
                                           ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double, _i64,
                                           ctypes.c_double, _D, _D, _D, _D, _D]
        L.orc_xcorr_image.argtypes = [_D, _D, _i64, _i64, _i64, ctypes.c_double, _D, _D, _D]
        L.orc_eikonal2d.argtypes = [_D, _i64, _i64, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, _D]
        L.orc_kirchhoff_migrate.argtypes = [_D, _i64, _i64, _i64, _I, ctypes.c_int, _I,
                                            ctypes.c_int, _D, _i64, ctypes.c_double,
                                            ctypes.c_int, _D]
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_freq_cpml_2d_nnz.restype = _i64
        L.orc_freq_cpml_2d_nnz.argtypes = [_i64, _i64, ctypes.c_int]
        _P64 = ctypes.POINTER(ctypes.c_int64)
        L.orc_freq_cpml_2d_matrix.argtypes = [_D, _i64, _i64, ctypes.c_double, ctypes.c_int,
                                              ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                              _P64, _P64, _D]
        L.orc_freq_cpml_2d_solve.argtypes = [_D, _i64, _i64, ctypes.c_double, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_double, ctypes.c_double, _D,
                                             ctypes.c_int, _D]
        _lib = L
    return _lib


def _dp(a):
    return None if a is None else a.ctypes.data_as(_D)


def _ip(a):
    return None if a is None else a.ctypes.data_as(_I)


def _f64(a):
    return np.asfortranarray(a, dtype=np.float64)


def _check(rc, name):
    if rc != 0:
        raise ValueError(f"oracle {name} failed with status {rc}")


def set_threads(n):
    lib().orc_set_threads(int(n))


def num_threads():
    return lib().orc_num_threads()


def dcoef(N):
    c = np.zeros(N)
    _check(lib().orc_dcoef(N, _dp(c)), "dcoef")
    return c


def cpml_profiles_2d(v, nB, dz, dx, dt):
    v = _f64(v)
    nz, nx = v.shape
    xb = np.zeros((nz, nx), order="F")
    zb = np.zeros((nz, nx), order="F")
    _check(lib().orc_cpml_profiles_2d(_dp(v), nz, nx, nB, dz, dx, dt, _dp(xb), _dp(zb)),
           "cpml_profiles_2d")
    return xb, zb


def cpml_profiles_3d(v, nB, dt):
    v = _f64(v)
    nz, nx, ny = v.shape
    xb, yb, zb = (np.zeros((nz, nx, ny), order="F") for _ in range(3))
    _check(lib().orc_cpml_profiles_3d(_dp(v), nz, nx, ny, nB, dt, _dp(xb), _dp(yb), _dp(zb)),
           "cpml_profiles_3d")
    return xb, yb, zb


def _points(src_iz, src_ix, wavelet, nt, src_iy=None):
    iz = np.ascontiguousarray(np.atleast_1d(src_iz), dtype=np.int32)
    ix = np.ascontiguousarray(np.atleast_1d(src_ix), dtype=np.int32)
    iy = None if src_iy is None else np.ascontiguousarray(np.atleast_1d(src_iy), dtype=np.int32)
    w = np.asarray(wavelet, dtype=np.float64)
    if w.ndim == 1:
        w = w[:, None]
    w = np.ascontiguousarray(np.broadcast_to(w, (nt, iz.size)))
    return iz, ix, iy, w


def _nsnap(nt, stride):
    return (nt + stride - 1) // stride


def fwd_cpml_2d(v, nt, N, nB, dz, dx, dt, source=None, src_iz=None, src_ix=None,
                wavelet=None, want_snapshot=False, snap_stride=1):
    """fwdTimeCpmlFor2dAw: returns (data[nx, nt], snapshot[nz, nx, nsnap] or None)."""
    v = _f64(v)
    nz, nx = v.shape
    data = np.zeros((nx, nt), order="F")
    snap = np.zeros((nz, nx, _nsnap(nt, snap_stride)), order="F") if want_snapshot else None
    if source is not None:
        src = _f64(source)
        assert src.shape == (nz, nx, nt)
        rc = lib().orc_fwd_cpml_2d(_dp(v), nz, nx, nt, _dp(src), 0, None, None, None, N, nB,
                                   dz, dx, dt, _dp(data), _dp(snap), snap_stride)
    else:
        iz, ix, _, w = _points(src_iz, src_ix, wavelet, nt)
        rc = lib().orc_fwd_cpml_2d(_dp(v), nz, nx, nt, None, iz.size, _ip(iz), _ip(ix), _dp(w),
                                   N, nB, dz, dx, dt, _dp(data), _dp(snap), snap_stride)
    _check(rc, "fwd_cpml_2d")
    return data, snap


def rvs_cpml_2d(v, data, N, nB, dz, dx, dt, want_snapshot=False, snap_stride=1):
    """rvsTimeCpmlFor2dAw: returns (model[nz+2l, nx+2l], snapshot or None)."""
    v = _f64(v)
    nz, nx = v.shape
    data = _f64(data)
    nt = data.shape[1]
    l = 2 * N - 1
    model = np.zeros((nz + 2 * l, nx + 2 * l), order="F")
    snap = np.zeros((nz, nx, _nsnap(nt, snap_stride)), order="F") if want_snapshot else None
    _check(lib().orc_rvs_cpml_2d(_dp(v), nz, nx, nt, _dp(data), N, nB, dz, dx, dt, _dp(model),
                                 _dp(snap), snap_stride), "rvs_cpml_2d")
    return model, snap


def fwd_cpml_3d(v, nt, N, nB, dz, dx, dy, dt, source=None, src_iz=None, src_ix=None,
                src_iy=None, wavelet=None, want_snapshot=False, snap_stride=1):
    """fwdTimeCpmlFor3dAw: returns (data[nx, ny, nt], snapshot[nz, nx, ny, nsnap] or None)."""
    v = _f64(v)
    nz, nx, ny = v.shape
    data = np.zeros((nx, ny, nt), order="F")
    snap = (np.zeros((nz, nx, ny, _nsnap(nt, snap_stride)), order="F")
            if want_snapshot else None)
    if source is not None:
        src = _f64(source)
        rc = lib().orc_fwd_cpml_3d(_dp(v), nz, nx, ny, nt, _dp(src), 0, None, None, None, None,
                                   N, nB, dz, dx, dy, dt, _dp(data), _dp(snap), snap_stride)
    else:
        iz, ix, iy, w = _points(src_iz, src_ix, wavelet, nt, src_iy)
        rc = lib().orc_fwd_cpml_3d(_dp(v), nz, nx, ny, nt, None, iz.size, _ip(iz), _ip(ix),
                                   _ip(iy), _dp(w), N, nB, dz, dx, dy, dt, _dp(data),
                                   _dp(snap), snap_stride)
    _check(rc, "fwd_cpml_3d")
    return data, snap


def fwd_sponge_2d(v, nt, nB, dz, dx, dt, source=None, src_iz=None, src_ix=None, wavelet=None,
                  want_snapshot=False, snap_stride=1):
    """fwdTimeSpongeFor2dAw: returns (data[nx, nt], snapshot or None)."""
    v = _f64(v)
    nz, nx = v.shape
    data = np.zeros((nx, nt), order="F")
    snap = np.zeros((nz, nx, _nsnap(nt, snap_stride)), order="F") if want_snapshot else None
    if source is not None:
        src = _f64(source)
        rc = lib().orc_fwd_sponge_2d(_dp(v), nz, nx, nt, _dp(src), 0, None, None, None, nB,
                                     dz, dx, dt, _dp(data), _dp(snap), snap_stride)
    else:
        iz, ix, _, w = _points(src_iz, src_ix, wavelet, nt)
        rc = lib().orc_fwd_sponge_2d(_dp(v), nz, nx, nt, None, iz.size, _ip(iz), _ip(ix),
                                     _dp(w), nB, dz, dx, dt, _dp(data), _dp(snap), snap_stride)
    _check(rc, "fwd_sponge_2d")
    return data, snap


def rvs_sponge_2d(v, data, nB, dz, dx, dt, want_snapshot=False, snap_stride=1):
    """rvsTimeSpongeFor2dAw: returns (model[nz+2, nx+2], snapshot or None)."""
    v = _f64(v)
    nz, nx = v.shape
    data = _f64(data)
    nt = data.shape[1]
    model = np.zeros((nz + 2, nx + 2), order="F")
    snap = np.zeros((nz, nx, _nsnap(nt, snap_stride)), order="F") if want_snapshot else None
    _check(lib().orc_rvs_sponge_2d(_dp(v), nz, nx, nt, _dp(data), nB, dz, dx, dt, _dp(model),
                                   _dp(snap), snap_stride), "rvs_sponge_2d")
    return model, snap


def rtm_shot_cpml_2d(V, VS, nt, src_iz, src_ix, wavelet, N, nB, dz, dx, dt, img_stride=1,
                     eps=1e-3, data_true=None):
    """One RTM shot: returns dict(num, den, image, data_true) (each [nz, nx] / [nx, nt])."""
    V = _f64(V)
    VS = _f64(VS)
    nz, nx = V.shape
    iz, ix, _, w = _points(src_iz, src_ix, wavelet, nt)
    dtin = None if data_true is None else _f64(data_true)
    dtout = np.zeros((nx, nt), order="F")
    num, den, img = (np.zeros((nz, nx), order="F") for _ in range(3))
    _check(lib().orc_rtm_shot_cpml_2d(_dp(V), _dp(VS), nz, nx, nt, iz.size, _ip(iz), _ip(ix),
                                      _dp(w), N, nB, dz, dx, dt, img_stride, eps, _dp(dtin),
                                      _dp(dtout), _dp(num), _dp(den), _dp(img)),
           "rtm_shot_cpml_2d")
    return {"num": num, "den": den, "image": img, "data_true": dtout}


def xcorr_image(fwd_snap, rvs_snap, stride=1, eps=1e-3):
    f = _f64(fwd_snap)
    r = _f64(rvs_snap)
    shp = f.shape[:-1]
    nt = f.shape[-1]
    ncell = int(np.prod(shp))
    num, den, img = (np.zeros(shp, order="F") for _ in range(3))
    _check(lib().orc_xcorr_image(_dp(f), _dp(r), ncell, nt, stride, eps, _dp(num), _dp(den),
                                 _dp(img)), "xcorr_image")
    return num, den, img


def eikonal2d(V, dx, sz, sx, iMax):
    """T = eikonal2d(V, dx, sz, sx, iMax) (src/fd_cmex/eikonal2d_mex.cpp); sz, sx 1-based."""
    V = _f64(V)
    nz, nx = V.shape
    T = np.zeros((nz, nx), order="F")
    _check(lib().orc_eikonal2d(_dp(V), nz, nx, float(dx), int(sz), int(sx), int(iMax), _dp(T)),
           "eikonal2d")
    return T


def kirchhoff_migrate(travelTime, xRecGrid, xShotAndRecGrid, shot, dt, nz, ixs):
    """m = migrate(travelTime, xRecGrid, xShotAndRecGrid, shot, dt, nz, ixs)
    (src/migrate.m); grids and ixs 1-based, shot (nt, nx)."""
    tt = _f64(travelTime)
    if tt.ndim == 2:
        tt = tt[:, :, None]
    nzt, nx, ntab = tt.shape
    assert nzt == nz
    shot = _f64(shot)
    xr = np.ascontiguousarray(np.atleast_1d(xRecGrid), dtype=np.int32)
    xs = np.ascontiguousarray(np.atleast_1d(xShotAndRecGrid), dtype=np.int32)
    m = np.zeros((nz, nx), order="F")
    _check(lib().orc_kirchhoff_migrate(_dp(tt), nz, nx, ntab, _ip(xr), xr.size, _ip(xs), xs.size,
                                       _dp(shot), shot.shape[0], float(dt), int(ixs), _dp(m)),
           "kirchhoff_migrate")
    return m


def freq_cpml_2d_matrix(model, w, N, nB, dz, dx):
    """A of freqCpmlFor2dAw (src/freqCpmlFor2dAw.m:43-123) as CSR arrays
    (rowptr, col, val complex128); model is the squared slowness (nz, nx)."""
    m = _f64(model)
    nz, nx = m.shape
    nnz = lib().orc_freq_cpml_2d_nnz(nz, nx, int(N))
    rp = np.zeros(nz * nx + 1, dtype=np.int64)
    col = np.zeros(nnz, dtype=np.int64)
    val = np.zeros(nnz, dtype=np.complex128)
    P64 = ctypes.POINTER(ctypes.c_int64)
    _check(lib().orc_freq_cpml_2d_matrix(_dp(m), nz, nx, float(w), int(N), int(nB), float(dz),
                                         float(dx), rp.ctypes.data_as(P64),
                                         col.ctypes.data_as(P64),
                                         val.view(np.float64).ctypes.data_as(_D)),
           "freq_cpml_2d_matrix")
    return rp, col, val


def freq_cpml_2d_solve(model, source, w, N, nB, dz, dx):
    """U = A \\ (-S) (src/freqCpmlFor2dAw.m:127-133) by banded LU with partial
    pivoting; source (nz*nx, nsrc) or (nz, nx[, nsrc]) complex."""
    m = _f64(model)
    nz, nx = m.shape
    S = np.asarray(source, dtype=np.complex128).reshape(nz * nx, -1, order="F")
    ns = S.shape[1]
    s1 = np.ascontiguousarray(S.ravel(order="F"))
    u1 = np.zeros(nz * nx * ns, dtype=np.complex128)
    _check(lib().orc_freq_cpml_2d_solve(_dp(m), nz, nx, float(w), int(N), int(nB), float(dz),
                                        float(dx), s1.view(np.float64).ctypes.data_as(_D), ns,
                                        u1.view(np.float64).ctypes.data_as(_D)),
           "freq_cpml_2d_solve")
    return u1.reshape(nz * nx, ns, order="F")

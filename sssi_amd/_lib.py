"""ctypes bindings of libsssi.so (include/sssi.h) and thin numpy wrappers.

Host arrays cross the boundary as column-major FP64, exactly as a MEX gateway
passes mxGetPr pointers; the library converts to the device layout.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsssi.so")

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_i64 = ctypes.c_int64
_int = ctypes.c_int
_dbl = ctypes.c_double
_vp = ctypes.c_void_p

# name -> argtypes (restype int unless listed in _RESTYPE)
SIGNATURES = {
    "sssi_last_error": [],
    "sssi_version": [],
    "sssi_device_count": [_I],
    "sssi_set_device": [_int],
    "sssi_set_autotune": [_int],
    "sssi_fwdTimeCpmlFor2dAw": [_D, _i64, _i64, _D, _i64, _int, _int, _dbl, _dbl, _dbl, _D, _D],
    "sssi_rvsTimeCpmlFor2dAw": [_D, _i64, _i64, _D, _i64, _int, _int, _dbl, _dbl, _dbl, _D, _D],
    "sssi_fwdTimeCpmlFor3dAw": [_D, _i64, _i64, _i64, _D, _i64, _int, _int, _dbl, _dbl, _dbl,
                                _dbl, _D, _D],
    "sssi_fwdTimeSpongeFor2dAw": [_D, _i64, _i64, _D, _i64, _int, _dbl, _dbl, _dbl, _D, _D],
    "sssi_rvsTimeSpongeFor2dAw": [_D, _i64, _i64, _D, _i64, _int, _dbl, _dbl, _dbl, _D, _D],
    "sssi_fwd_cpml_2d_points": [_D, _i64, _i64, _i64, _int, _I, _I, _D, _int, _int, _dbl, _dbl,
                                _dbl, _D, _D, _i64],
    "sssi_fwd_cpml_2d_points_noalloc": [_D, _i64, _i64, _i64, _int, _I, _I, _D, _int, _int, _dbl,
                                        _dbl, _dbl, _D, _D, _i64],
    "sssi_rvs_cpml_2d_ex": [_D, _i64, _i64, _D, _i64, _int, _int, _dbl, _dbl, _dbl, _D, _D, _i64],
    "sssi_fwd_cpml_3d_points": [_D, _i64, _i64, _i64, _i64, _int, _I, _I, _I, _D, _int, _int,
                                _dbl, _dbl, _dbl, _dbl, _D, _D, _i64],
    "sssi_fwd_sponge_2d_points": [_D, _i64, _i64, _i64, _int, _I, _I, _D, _int, _dbl, _dbl,
                                  _dbl, _D, _D, _i64],
    "sssi_rvs_sponge_2d_ex": [_D, _i64, _i64, _D, _i64, _int, _dbl, _dbl, _dbl, _D, _D, _i64],
    "sssi_prop2d_create": [_D, _i64, _i64, _int, _int, _int, _dbl, _dbl, _dbl,
                           ctypes.POINTER(_vp)],
    "sssi_prop3d_create": [_D, _i64, _i64, _i64, _int, _int, _dbl, _dbl, _dbl, _dbl,
                           ctypes.POINTER(_vp)],
    "sssi_prop_destroy": [_vp],
    "sssi_prop_set_points": [_vp, _int, _I, _I, _I, _D, _i64],
    "sssi_prop_set_surface": [_vp, _D, _i64],
    "sssi_prop_reset": [_vp],
    "sssi_prop_run": [_vp, _i64, _int, _int],
    "sssi_prop_run_steps": [_vp, _i64, _i64, _int, _int],
    "sssi_prop_run_timed": [_vp, _i64, _int, _int, _D, ctypes.POINTER(_i64)],
    "sssi_prop_sync": [_vp],
    "sssi_prop_get_data": [_vp, _D, _i64],
    "sssi_prop_get_field": [_vp, _D],
    "sssi_prop_get_model": [_vp, _D],
    "sssi_prop_alg_bytes_per_step": [_vp, _D, _D],
    "sssi_prop_layout": [_vp, _I, _I],
    "sssi_rtm2d_create": [_D, _D, _i64, _i64, _int, _int, _dbl, _dbl, _dbl, _i64, _int, _int, _dbl,
                          ctypes.POINTER(_vp)],
    "sssi_rtm2d_add_shot": [_vp, _i64, _i64, _int, _D, _I],
    "sssi_rtm2d_precompute_data": [_vp, _int],
    "sssi_rtm2d_migrate": [_vp, _int, _D, ctypes.POINTER(_i64)],
    "sssi_rtm2d_get_stack": [_vp, _D],
    "sssi_rtm2d_get_shot": [_vp, _int, _D, _D, _D, _D],
    "sssi_rtm2d_set_cache": [_vp, ctypes.c_char_p],
    "sssi_rtm2d_cache_stats": [_vp, _I, _I, _I],
    "sssi_rtm2d_device_bytes": [_vp, _D],
    "sssi_eikonal2d": [_D, _i64, _i64, _dbl, _int, _int, _int, _D],
    "sssi_eikonal2d_batch": [_D, _i64, _i64, _dbl, _int, _I, _I, _int, _D],
    "sssi_kirchhoff_migrate": [_D, _i64, _i64, _i64, _I, _int, _I, _int, _D, _i64, _dbl, _int,
                               _D],
    "sssi_kirchhoff_create": [_D, _i64, _i64, _dbl, _I, _int, _int, _dbl, ctypes.POINTER(_vp)],
    "sssi_kirchhoff_travel_time": [_vp, _D],
    "sssi_kirchhoff_migrate_shots": [_vp, _I, _int, _D, _i64, _int, _I, _D, _D],
    "sssi_kirchhoff_load_shots": [_vp, _D, _i64, _int],
    "sssi_kirchhoff_migrate_loaded": [_vp, _I, _int, _I, _D, _D],
    "sssi_kirchhoff_timing": [_vp, _D, ctypes.POINTER(_i64), ctypes.POINTER(_i64), _D],
    "sssi_kirchhoff_destroy": [_vp],
    "sssi_calib_stream": [_i64, _int, _D, _D],
    "sssi_freqCpmlFor2dAw": [_D, _i64, _i64, _D, _int, _dbl, _int, _int, _dbl, _dbl, _D],
    "sssi_freq_cpml_2d_matrix": [_D, _i64, _i64, _dbl, _int, _int, _dbl, _dbl,
                                 ctypes.POINTER(_i64), ctypes.POINTER(_i64),
                                 ctypes.POINTER(_i64), _D],
    "sssi_freq_last_timing": [_D, _D],
    "sssi_mat_write": [ctypes.c_char_p, _int, ctypes.POINTER(ctypes.c_char_p),
                       ctypes.POINTER(_D), _I, ctypes.POINTER(_i64)],
    "sssi_mat_read": [ctypes.c_char_p, ctypes.c_char_p, _I, ctypes.POINTER(_i64), _D, _i64],
    "sssi_rtm2d_destroy": [_vp],
    "sssi_server_create": [_dbl, _int, ctypes.POINTER(_vp)],
    "sssi_server_submit_fwd2d": [_vp, _D, _i64, _i64, _i64, _int, _int, _dbl, _dbl, _dbl, _int,
                                 _int, _D, ctypes.POINTER(_i64)],
    "sssi_server_submit_rtm2d": [_vp, _D, _D, _i64, _i64, _i64, _int, _int, _dbl, _dbl, _dbl,
                                 _int, _D, _int, _int, _dbl, ctypes.POINTER(_i64)],
    "sssi_server_submit_rtm2d_data": [_vp, _D, _D, _i64, _i64, _i64, _int, _int, _dbl, _dbl, _dbl,
                                      _int, _D, _D, _int, _int, _dbl, ctypes.POINTER(_i64)],
    "sssi_server_submit_rtm2d_borrowed": [_vp, _D, _D, _i64, _i64, _i64, _int, _int, _dbl, _dbl,
                                          _dbl, _int, _D, _D, _int, _int, _dbl,
                                          ctypes.POINTER(_i64)],
    "sssi_server_submit_fwd3d": [_vp, _D, _i64, _i64, _i64, _i64, _int, _int, _dbl, _dbl, _dbl,
                                 _dbl, _int, _int, _int, _D, ctypes.POINTER(_i64)],
    "sssi_server_wait": [_vp, _i64, _dbl, _I],
    "sssi_server_drain": [_vp],
    "sssi_server_result": [_vp, _i64, _D, _i64],
    "sssi_server_job_info": [_vp, _i64, _D, _D, _D, _D],
    "sssi_server_stats": [_vp, _D, _D, ctypes.POINTER(_i64), ctypes.POINTER(_i64)],
    "sssi_server_destroy": [_vp],
}
_RESTYPE = {"sssi_last_error": ctypes.c_char_p, "sssi_prop_destroy": None,
            "sssi_rtm2d_destroy": None, "sssi_server_destroy": None,
            "sssi_kirchhoff_destroy": None}

_lib = None


class SssiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"sssi status {code}: {msg}")
        self.code = code


def lib():
    """Load libsssi.so (built by sssi_amd.build / __graft_entry__.build)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m sssi_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RESTYPE.get(name, ctypes.c_int)
        _lib = L
    return _lib


def _call(name, *args):
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        msg = lib().sssi_last_error().decode(errors="replace")
        raise SssiError(rc, msg)


def device_count():
    n = ctypes.c_int(0)
    _call("sssi_device_count", ctypes.byref(n))
    return n.value


def set_autotune(on=True):
    """Handles created afterwards time the kernel layouts on their own grid and
    keep the fastest (cached per problem shape); see sssi_set_autotune."""
    _call("sssi_set_autotune", 1 if on else 0)


def _dp(a):
    return None if a is None else a.ctypes.data_as(_D)


def _ip(a):
    return None if a is None else a.ctypes.data_as(_I)


def _f64(a):
    return np.asfortranarray(a, dtype=np.float64)


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


def fwd_cpml_2d_points(v, nt, nDiffOrder, nBoundary, dz, dx, dt, src_iz, src_ix, wavelet,
                       want_snapshot=False, snap_stride=1):
    """2-D CPML forward modelling with point sources; returns (data[nx,nt], snapshot|None)."""
    v = _f64(v)
    nz, nx = v.shape
    iz, ix, _, w = _points(src_iz, src_ix, wavelet, nt)
    data = np.zeros((nx, nt), order="F")
    snap = np.zeros((nz, nx, _nsnap(nt, snap_stride)), order="F") if want_snapshot else None
    _call("sssi_fwd_cpml_2d_points", _dp(v), nz, nx, nt, iz.size, _ip(iz), _ip(ix), _dp(w),
          int(nDiffOrder), int(nBoundary), float(dz), float(dx), float(dt), _dp(data), _dp(snap),
          snap_stride)
    return data, snap


def rvs_cpml_2d_ex(v, data, nDiffOrder, nBoundary, dz, dx, dt, want_snapshot=True,
                   snap_stride=1):
    v = _f64(v)
    nz, nx = v.shape
    data = _f64(data)
    nt = data.shape[1]
    l = 2 * int(nDiffOrder) - 1
    model = np.zeros((nz + 2 * l, nx + 2 * l), order="F")
    snap = np.zeros((nz, nx, _nsnap(nt, snap_stride)), order="F") if want_snapshot else None
    _call("sssi_rvs_cpml_2d_ex", _dp(v), nz, nx, _dp(data), nt, int(nDiffOrder), int(nBoundary),
          float(dz), float(dx), float(dt), _dp(model), _dp(snap), snap_stride)
    return model, snap


def fwd_cpml_3d_points(v, nt, nDiffOrder, nBoundary, dz, dx, dy, dt, src_iz, src_ix, src_iy,
                       wavelet, want_snapshot=False, snap_stride=1):
    v = _f64(v)
    nz, nx, ny = v.shape
    iz, ix, iy, w = _points(src_iz, src_ix, wavelet, nt, src_iy)
    data = np.zeros((nx, ny, nt), order="F")
    snap = (np.zeros((nz, nx, ny, _nsnap(nt, snap_stride)), order="F")
            if want_snapshot else None)
    _call("sssi_fwd_cpml_3d_points", _dp(v), nz, nx, ny, nt, iz.size, _ip(iz), _ip(ix), _ip(iy),
          _dp(w), int(nDiffOrder), int(nBoundary), float(dz), float(dx), float(dy), float(dt),
          _dp(data), _dp(snap), snap_stride)
    return data, snap


def fwd_sponge_2d_points(v, nt, nBoundary, dz, dx, dt, src_iz, src_ix, wavelet,
                         want_snapshot=False, snap_stride=1):
    v = _f64(v)
    nz, nx = v.shape
    iz, ix, _, w = _points(src_iz, src_ix, wavelet, nt)
    data = np.zeros((nx, nt), order="F")
    snap = np.zeros((nz, nx, _nsnap(nt, snap_stride)), order="F") if want_snapshot else None
    _call("sssi_fwd_sponge_2d_points", _dp(v), nz, nx, nt, iz.size, _ip(iz), _ip(ix), _dp(w),
          int(nBoundary), float(dz), float(dx), float(dt), _dp(data), _dp(snap), snap_stride)
    return data, snap


def rvs_sponge_2d_ex(v, data, nBoundary, dz, dx, dt, want_snapshot=True, snap_stride=1):
    v = _f64(v)
    nz, nx = v.shape
    data = _f64(data)
    nt = data.shape[1]
    model = np.zeros((nz + 2, nx + 2), order="F")
    snap = np.zeros((nz, nx, _nsnap(nt, snap_stride)), order="F") if want_snapshot else None
    _call("sssi_rvs_sponge_2d_ex", _dp(v), nz, nx, _dp(data), nt, int(nBoundary), float(dz),
          float(dx), float(dt), _dp(model), _dp(snap), snap_stride)
    return model, snap


class Propagator:
    """Device-resident propagation handle (sssi_prop): the model, damping
    profiles, sources and wavefields stay in HBM across runs."""

    def __init__(self, v, nDiffOrder, nBoundary, dz, dx, dt, dy=None, abc="cpml"):
        v = _f64(v)
        self.shape = v.shape
        h = ctypes.c_void_p()
        if v.ndim == 2:
            code = 0 if abc == "cpml" else 1
            _call("sssi_prop2d_create", _dp(v), v.shape[0], v.shape[1], code, int(nDiffOrder),
                  int(nBoundary), float(dz), float(dx), float(dt), ctypes.byref(h))
        else:
            _call("sssi_prop3d_create", _dp(v), v.shape[0], v.shape[1], v.shape[2],
                  int(nDiffOrder), int(nBoundary), float(dz), float(dx), float(dy), float(dt),
                  ctypes.byref(h))
        self._h = h
        self.nDiffOrder = int(nDiffOrder)
        self.abc = abc

    def close(self):
        if self._h:
            lib().sssi_prop_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_points(self, src_iz, src_ix, wavelet, src_iy=None):
        w = np.asarray(wavelet, dtype=np.float64)
        nt = w.shape[0]
        iz, ix, iy, w = _points(src_iz, src_ix, w, nt, src_iy)
        _call("sssi_prop_set_points", self._h, iz.size, _ip(iz), _ip(ix), _ip(iy), _dp(w), nt)

    def set_surface(self, data):
        d = _f64(data)
        _call("sssi_prop_set_surface", self._h, _dp(d), d.shape[1])

    def reset(self):
        _call("sssi_prop_reset", self._h)

    def run(self, nt, reverse=False, record=True):
        _call("sssi_prop_run", self._h, int(nt), int(bool(reverse)), int(bool(record)))

    def run_steps(self, t_first, count, reverse=False, record=True):
        """Continue from time index t_first (see sssi_prop_run_steps)."""
        _call("sssi_prop_run_steps", self._h, int(t_first), int(count), int(bool(reverse)),
              int(bool(record)))

    def run_timed(self, nt, reverse=False, record=True):
        ms = ctypes.c_double(0)
        n = ctypes.c_int64(0)
        _call("sssi_prop_run_timed", self._h, int(nt), int(bool(reverse)), int(bool(record)),
              ctypes.byref(ms), ctypes.byref(n))
        return ms.value, n.value

    def sync(self):
        _call("sssi_prop_sync", self._h)

    def data(self, nt):
        if len(self.shape) == 2:
            out = np.zeros((self.shape[1], nt), order="F")
        else:
            out = np.zeros((self.shape[1], self.shape[2], nt), order="F")
        _call("sssi_prop_get_data", self._h, _dp(out), int(nt))
        return out

    def field(self):
        out = np.zeros(self.shape, order="F")
        _call("sssi_prop_get_field", self._h, _dp(out))
        return out

    def model(self):
        l = 2 * self.nDiffOrder - 1 if self.abc == "cpml" else 1
        out = np.zeros((self.shape[0] + 2 * l, self.shape[1] + 2 * l), order="F")
        _call("sssi_prop_get_model", self._h, _dp(out))
        return out

    def alg_bytes_per_step(self):
        b = ctypes.c_double(0)
        c = ctypes.c_double(0)
        _call("sssi_prop_alg_bytes_per_step", self._h, ctypes.byref(b), ctypes.byref(c))
        return b.value, c.value

    def layout(self):
        """2-D: (tile height, interior march groups); 3-D: (block y extent, z chunk)."""
        a = ctypes.c_int(0)
        b = ctypes.c_int(0)
        _call("sssi_prop_layout", self._h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

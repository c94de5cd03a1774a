"""Kirchhoff path over the C ABI: traveltimes by the reference's fast-sweeping
eikonal solver (src/fd_cmex/eikonal2d_mex.cpp) and Kirchhoff migration
(src/migrate.m), same names, argument order and 1-based positions as the
MATLAB functions; a resident handle for the app's loops
(applications/kirchhoff_2d_acoustic/mainKirchhoffCpmlFor2dAw.m:244-293)."""
import ctypes

import numpy as np

from ._lib import _call, _dp, _f64, _ip, lib


def _ints(a):
    return np.ascontiguousarray(np.atleast_1d(a), dtype=np.int32)


def eikonal2d(V, dx, sz, sx, iMax):
    """T = eikonal2d(V, dx, sz, sx, iMax); sz, sx 1-based; T (nz, nx)."""
    V = _f64(V)
    nz, nx = V.shape
    T = np.zeros((nz, nx), order="F")
    _call("sssi_eikonal2d", _dp(V), nz, nx, float(dx), int(sz), int(sx), int(iMax), _dp(T))
    return T


def eikonal2d_batch(V, dx, sz, sx, iMax):
    """Traveltimes of several sources at once: T (nz, nx, nsrc)."""
    V = _f64(V)
    nz, nx = V.shape
    z, x = _ints(sz), _ints(sx)
    T = np.zeros((nz, nx, z.size), order="F")
    _call("sssi_eikonal2d_batch", _dp(V), nz, nx, float(dx), z.size, _ip(z), _ip(x), int(iMax),
          _dp(T))
    return T


def migrate(travelTime, xRecGrid, xShotAndRecGrid, shot, dt, nz, ixs):
    """m = migrate(travelTime, xRecGrid, xShotAndRecGrid, shot, dt, nz, ixs)."""
    tt = _f64(travelTime)
    if tt.ndim == 2:
        tt = tt[:, :, None]
    if tt.shape[0] != nz:
        raise ValueError("travelTime must have nz rows")
    shot = _f64(shot)
    xr, xs = _ints(xRecGrid), _ints(xShotAndRecGrid)
    m = np.zeros((nz, tt.shape[1]), order="F")
    _call("sssi_kirchhoff_migrate", _dp(tt), nz, tt.shape[1], tt.shape[2], _ip(xr), xr.size,
          _ip(xs), xs.size, _dp(shot), shot.shape[0], float(dt), int(ixs), _dp(m))
    return m


class Kirchhoff2D:
    """Traveltime tables of the surface positions xShotAndRecGrid (sources at
    row 1) on the device, and the migration of many shots against them."""

    def __init__(self, V, dx, xShotAndRecGrid, iMax, dt):
        V = _f64(V)
        self.nz, self.nx = V.shape
        self.grid = _ints(xShotAndRecGrid)
        h = ctypes.c_void_p()
        _call("sssi_kirchhoff_create", _dp(V), self.nz, self.nx, float(dx), _ip(self.grid),
              self.grid.size, int(iMax), float(dt), ctypes.byref(h))
        self._h = h

    def close(self):
        if self._h:
            lib().sssi_kirchhoff_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def travel_time(self):
        T = np.zeros((self.nz, self.nx, self.grid.size), order="F")
        _call("sssi_kirchhoff_travel_time", self._h, _dp(T))
        return T

    def timing(self):
        """(eikonal kernel ms, passes, launches, last migration kernel ms), from
        HIP events on the handle's stream."""
        e, m = ctypes.c_double(), ctypes.c_double()
        n, nl = ctypes.c_int64(), ctypes.c_int64()
        _call("sssi_kirchhoff_timing", self._h, ctypes.byref(e), ctypes.byref(n),
              ctypes.byref(nl), ctypes.byref(m))
        return e.value, n.value, nl.value, m.value

    def load_shots(self, shots):
        """Keep records (nt, nx, nshot) on the device for migrate_loaded."""
        s = _f64(shots)
        if s.ndim == 2:
            s = s[:, :, None]
        _call("sssi_kirchhoff_load_shots", self._h, _dp(s), s.shape[0], s.shape[2])
        self._loaded = s.shape[2]

    def migrate_loaded(self, xRecGrid, ixs, images=False):
        """Migrate the loaded records: stack (nz, nx) [, images (nz, nx, nshot)]."""
        ix = _ints(ixs)
        if ix.size != getattr(self, "_loaded", -1):
            raise ValueError("one table index per loaded shot")
        xr = _ints(xRecGrid)
        st = np.zeros((self.nz, self.nx), order="F")
        im = np.zeros((self.nz, self.nx, ix.size), order="F") if images else None
        _call("sssi_kirchhoff_migrate_loaded", self._h, _ip(xr), xr.size, _ip(ix), _dp(st),
              _dp(im))
        return (st, im) if images else st

    def migrate_shots(self, xRecGrid, shots, ixs, images=False):
        """shots (nt, nx, nshot) with table indices ixs (1-based): the stack
        (nz, nx) and, if asked, the per-shot images (nz, nx, nshot)."""
        s = _f64(shots)
        if s.ndim == 2:
            s = s[:, :, None]
        ix = _ints(ixs)
        if ix.size != s.shape[2]:
            raise ValueError("one table index per shot")
        xr = _ints(xRecGrid)
        st = np.zeros((self.nz, self.nx), order="F")
        im = np.zeros((self.nz, self.nx, ix.size), order="F") if images else None
        _call("sssi_kirchhoff_migrate_shots", self._h, _ip(xr), xr.size, _dp(s), s.shape[0],
              ix.size, _ip(ix), _dp(st), _dp(im))
        self._loaded = ix.size
        return (st, im) if images else st

"""MATLAB-signature drop-ins over the C ABI (same names, argument order and
outputs as the reference wrappers; see the package docstring)."""
import numpy as np

from ._lib import _call, _dp, _f64


def _snap_or_none(nargout, shape):
    return np.zeros(shape, order="F") if nargout >= 2 else None


def fwdTimeCpmlFor2dAw(v, source, nDiffOrder, nBoundary, dz, dx, dt, nargout=2):
    """src/fd_cmex/fwdTimeCpmlFor2dAw.m:37 -> fwdTimeCpmlFor2dAw_mex.c:39.
    v(nz,nx), source(nz,nx,nt) -> data(nx,nt), snapshot(nz,nx,nt)."""
    v = _f64(v)
    src = _f64(source)
    if src.ndim == 2:
        src = src[:, :, None]
    nz, nx = v.shape
    if src.shape[:2] != (nz, nx):
        raise ValueError("Velocity model and source grids should have the same size")
    nt = src.shape[2]
    data = np.zeros((nx, nt), order="F")
    snap = _snap_or_none(nargout, (nz, nx, nt))
    _call("sssi_fwdTimeCpmlFor2dAw", _dp(v), nz, nx, _dp(src), nt, int(nDiffOrder),
          int(nBoundary), float(dz), float(dx), float(dt), _dp(data), _dp(snap))
    return data, snap


def rvsTimeCpmlFor2dAw(v, data, nDiffOrder, nBoundary, dz, dx, dt, nargout=2):
    """src/fd_cmex/rvsTimeCpmlFor2dAw.m:30 -> rvsTimeCpmlFor2dAw_mex.c:38.
    v(nz,nx), data(nx,nt) -> model(nz+2l,nx+2l), snapshot(nz,nx,nt)."""
    v = _f64(v)
    d = _f64(data)
    nz, nx = v.shape
    if d.shape[0] != nx:
        raise ValueError("Velocity model and input data should have the same x-axis grids")
    nt = d.shape[1]
    l = 2 * int(nDiffOrder) - 1
    model = np.zeros((nz + 2 * l, nx + 2 * l), order="F")
    snap = _snap_or_none(nargout, (nz, nx, nt))
    _call("sssi_rvsTimeCpmlFor2dAw", _dp(v), nz, nx, _dp(d), nt, int(nDiffOrder),
          int(nBoundary), float(dz), float(dx), float(dt), _dp(model), _dp(snap))
    return model, snap


def fwdTimeCpmlFor3dAw(v, source, nDiffOrder, nBoundary, dz, dx, dy, dt, nargout=2):
    """src/fwdTimeCpmlFor3dAw.m:1.  v(nz,nx,ny), source(nz,nx,ny,nt) ->
    data(nx,ny,nt), snapshot(nz,nx,ny,nt)."""
    v = _f64(v)
    src = _f64(source)
    nz, nx, ny = v.shape
    if src.shape[:3] != (nz, nx, ny):
        raise ValueError("Velocity model and source grids should have the same size")
    nt = src.shape[3]
    data = np.zeros((nx, ny, nt), order="F")
    snap = _snap_or_none(nargout, (nz, nx, ny, nt))
    _call("sssi_fwdTimeCpmlFor3dAw", _dp(v), nz, nx, ny, _dp(src), nt, int(nDiffOrder),
          int(nBoundary), float(dz), float(dx), float(dy), float(dt), _dp(data), _dp(snap))
    return data, snap


def fwdTimeSpongeFor2dAw(v, source, nBoundary, dz, dx, dt, nargout=2):
    """src/fwdTimeSpongeFor2dAw.m:1.  -> data(nx,nt), snapshot(nz,nx,nt)."""
    v = _f64(v)
    src = _f64(source)
    nz, nx = v.shape
    nt = src.shape[2]
    data = np.zeros((nx, nt), order="F")
    snap = _snap_or_none(nargout, (nz, nx, nt))
    _call("sssi_fwdTimeSpongeFor2dAw", _dp(v), nz, nx, _dp(src), nt, int(nBoundary), float(dz),
          float(dx), float(dt), _dp(data), _dp(snap))
    return data, snap


def rvsTimeSpongeFor2dAw(v, data, nBoundary, dz, dx, dt, nargout=2):
    """src/rvsTimeSpongeFor2dAw.m:1.  -> model(nz+2,nx+2), snapshot(nz,nx,nt)."""
    v = _f64(v)
    d = _f64(data)
    nz, nx = v.shape
    nt = d.shape[1]
    model = np.zeros((nz + 2, nx + 2), order="F")
    snap = _snap_or_none(nargout, (nz, nx, nt))
    _call("sssi_rvsTimeSpongeFor2dAw", _dp(v), nz, nx, _dp(d), nt, int(nBoundary), float(dz),
          float(dx), float(dt), _dp(model), _dp(snap))
    return model, snap

"""sssi_amd -- MI355X-native acoustic FD/CPML propagators and RTM (SSSI hot path).

Host-side mirror of the reference's MATLAB interface over the C ABI of
``libsssi.so`` (include/sssi.h).  Function names, argument order, argument
meaning and outputs follow the MATLAB wrappers they replace:

  data, snapshot = fwdTimeCpmlFor2dAw(v, source, nDiffOrder, nBoundary, dz, dx, dt)
  model, snapshot = rvsTimeCpmlFor2dAw(v, data, nDiffOrder, nBoundary, dz, dx, dt)
  data, snapshot = fwdTimeCpmlFor3dAw(v, source, nDiffOrder, nBoundary, dz, dx, dy, dt)
  data, snapshot = fwdTimeSpongeFor2dAw(v, source, nBoundary, dz, dx, dt)
  model, snapshot = rvsTimeSpongeFor2dAw(v, data, nBoundary, dz, dx, dt)
  T = eikonal2d(V, dx, sz, sx, iMax);  m = migrate(travelTime, xRecGrid, ...)
  A, snapshot = freqCpmlFor2dAw(model, source, w, nDiffOrder, nBoundary, dz, dx)

Arrays are numpy arrays indexed like the MATLAB ones (``v[iz, ix]``,
``source[iz, ix, it]``, ``data[ix, it]``, ``snapshot[iz, ix, it]``).  As with
MATLAB's ``nargout``, pass ``nargout=1`` to skip the snapshot stack.
Errors raise :class:`SssiError` carrying the ABI status and message.

The compute runs on the GPU through libsssi; there is no CPU fallback.  When
the library is missing or no device is visible, calls raise.
"""
from ._lib import (SssiError, lib, device_count, set_autotune, fwd_cpml_2d_points,
                   rvs_cpml_2d_ex, fwd_cpml_3d_points, fwd_sponge_2d_points, rvs_sponge_2d_ex,
                   Propagator)
from .api import (fwdTimeCpmlFor2dAw, rvsTimeCpmlFor2dAw, fwdTimeCpmlFor3dAw,
                  fwdTimeSpongeFor2dAw, rvsTimeSpongeFor2dAw)
from . import synth
from .rtm import Rtm2D
from .serve import Server
from .matfile import mat_read, mat_write
from .kirchhoff import eikonal2d, eikonal2d_batch, migrate, Kirchhoff2D
from .helmholtz import freqCpmlFor2dAw

__all__ = [
    "SssiError", "lib", "device_count", "fwdTimeCpmlFor2dAw", "rvsTimeCpmlFor2dAw",
    "fwdTimeCpmlFor3dAw", "fwdTimeSpongeFor2dAw", "rvsTimeSpongeFor2dAw",
    "fwd_cpml_2d_points", "rvs_cpml_2d_ex", "fwd_cpml_3d_points", "fwd_sponge_2d_points",
    "rvs_sponge_2d_ex", "Propagator", "synth", "Rtm2D", "Server", "mat_read", "mat_write",
    "eikonal2d", "eikonal2d_batch", "migrate", "Kirchhoff2D", "freqCpmlFor2dAw",
]

"""MAT-file (Level 5) I/O of real double arrays over the C ABI (sssi_mat_*).

The durable format of the per-shot RTM results (see Rtm2D.set_cache): MATLAB's
`load` reads these files, as the reference's apps load their per-shot
dataTrue<xs>.mat / rtmsnapshot<xs>.mat and StackedRTM.mat
(applications/rtm_2d_acoustic/mainRtmTimeCpmlFor2dAw.m:170-184, :247-264).
"""
import ctypes

import numpy as np

from ._lib import _call, _f64

_D = ctypes.POINTER(ctypes.c_double)


def mat_write(path, **variables):
    """Write the named arrays (any shape, stored column-major as doubles)."""
    if not variables:
        raise ValueError("nothing to write")
    names, arrays = [], []
    for k, v in variables.items():
        a = np.asarray(v, dtype=np.float64)
        if a.ndim < 2:
            a = a.reshape(-1, 1) if a.ndim == 1 else a.reshape(1, 1)
        names.append(k.encode())
        arrays.append(_f64(a))
    n = len(arrays)
    cnames = (ctypes.c_char_p * n)(*names)
    cdata = (_D * n)(*[a.ctypes.data_as(_D) for a in arrays])
    nd = (ctypes.c_int * n)(*[a.ndim for a in arrays])
    dims = [d for a in arrays for d in a.shape]
    cdims = (ctypes.c_int64 * len(dims))(*dims)
    _call("sssi_mat_write", str(path).encode(), n, cnames, cdata, nd, cdims)


def mat_read(path, name):
    """The variable `name` as a Fortran-ordered float64 array of its MAT shape."""
    nd = ctypes.c_int(0)
    dims = (ctypes.c_int64 * 8)()
    _call("sssi_mat_read", str(path).encode(), name.encode(), ctypes.byref(nd), dims, None, 0)
    shape = tuple(dims[k] for k in range(nd.value))
    out = np.zeros(shape, order="F")
    _call("sssi_mat_read", str(path).encode(), name.encode(), None, None,
          out.ctypes.data_as(_D), int(out.size))
    return out

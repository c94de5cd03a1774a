"""Serving loop with continuous batching over the C ABI (sssi_server_*).

Jobs are submitted at any time; the library's worker thread admits them by
slot and device-memory budget, advances every resident job one time step per
tick (2-D jobs of one nDiffOrder share a launch) and releases their memory when
their results are copied out.
"""
import ctypes

import numpy as np

from ._lib import _call, _dp, _f64, lib
from .rtm import CK_FORMATS


class Server:
    def __init__(self, hbm_budget_bytes=64e9, max_active=64):
        h = ctypes.c_void_p()
        _call("sssi_server_create", float(hbm_budget_bytes), int(max_active), ctypes.byref(h))
        self._h = h
        self._shapes = {}
        self._borrowed = {}  # job id -> arrays the library reads until the job is done

    def close(self):
        if self._h:
            lib().sssi_server_destroy(self._h)
            self._h = None
        self._borrowed.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit_fwd2d(self, v, nt, nDiffOrder, nBoundary, dz, dx, dt, src_iz, src_ix, wavelet):
        v = _f64(v)
        w = np.ascontiguousarray(wavelet, dtype=np.float64)
        jid = ctypes.c_int64(0)
        _call("sssi_server_submit_fwd2d", self._h, _dp(v), v.shape[0], v.shape[1], int(nt),
              int(nDiffOrder), int(nBoundary), float(dz), float(dx), float(dt), int(src_iz),
              int(src_ix), _dp(w), ctypes.byref(jid))
        self._shapes[jid.value] = (v.shape[1], int(nt))
        return jid.value

    def submit_rtm2d(self, V, VS, nt, nDiffOrder, nBoundary, dz, dx, dt, src_ix, wavelet,
                     img_stride=1, ck_format="fp32", eps=1e-3, data_true=None, borrow=True):
        """One RTM shot on the extended grid of V; with data_true (nx, nt), the
        cached observed data, the job skips its fwd(V) propagation.  With
        borrow (the default) the library reads V, VS and data_true when it sets
        the job up rather than copying them here; this object holds them until
        the job is done, and the caller must not modify them before then."""
        V = _f64(V)
        VS = _f64(VS)
        w = np.ascontiguousarray(wavelet, dtype=np.float64)
        dtr = None
        if data_true is not None:
            dtr = _f64(data_true)
            if dtr.shape != (V.shape[1], int(nt)):
                raise ValueError("data_true must be (nx, nt)")
        jid = ctypes.c_int64(0)
        entry = "sssi_server_submit_rtm2d_borrowed" if borrow else "sssi_server_submit_rtm2d_data"
        _call(entry, self._h, _dp(V), _dp(VS), V.shape[0], V.shape[1],
              int(nt), int(nDiffOrder), int(nBoundary), float(dz), float(dx), float(dt),
              int(src_ix), _dp(w), _dp(dtr), int(img_stride), CK_FORMATS[ck_format], float(eps),
              ctypes.byref(jid))
        self._shapes[jid.value] = V.shape
        if borrow:
            self._borrowed[jid.value] = (V, VS, dtr)
        return jid.value

    def submit_fwd3d(self, v, nt, nDiffOrder, nBoundary, dz, dx, dy, dt, src, wavelet):
        v = _f64(v)
        w = np.ascontiguousarray(wavelet, dtype=np.float64)
        jid = ctypes.c_int64(0)
        iz, ix, iy = src
        _call("sssi_server_submit_fwd3d", self._h, _dp(v), v.shape[0], v.shape[1], v.shape[2],
              int(nt), int(nDiffOrder), int(nBoundary), float(dz), float(dx), float(dy),
              float(dt), int(iz), int(ix), int(iy), _dp(w), ctypes.byref(jid))
        self._shapes[jid.value] = (v.shape[1], v.shape[2], int(nt))
        return jid.value

    def wait(self, job_id, timeout=-1.0):
        done = ctypes.c_int(0)
        _call("sssi_server_wait", self._h, int(job_id), float(timeout), ctypes.byref(done))
        if done.value:
            self._borrowed.pop(job_id, None)
        return bool(done.value)

    def drain(self):
        _call("sssi_server_drain", self._h)
        self._borrowed.clear()

    def result(self, job_id):
        shape = self._shapes[job_id]
        out = np.zeros(shape, order="F")
        _call("sssi_server_result", self._h, int(job_id), _dp(out), int(out.size))
        self._borrowed.pop(job_id, None)  # a job with a result is done
        return out

    def info(self, job_id):
        a, b, c, d = (ctypes.c_double(0) for _ in range(4))
        _call("sssi_server_job_info", self._h, int(job_id), ctypes.byref(a), ctypes.byref(b),
              ctypes.byref(c), ctypes.byref(d))
        return {"t_arrive": a.value, "t_admit": b.value, "t_done": c.value, "bytes": d.value,
                "latency": c.value - a.value}

    def stats(self):
        cu, busy = ctypes.c_double(0), ctypes.c_double(0)
        la, ti = ctypes.c_int64(0), ctypes.c_int64(0)
        _call("sssi_server_stats", self._h, ctypes.byref(cu), ctypes.byref(busy),
              ctypes.byref(la), ctypes.byref(ti))
        return {"cell_updates": cu.value, "busy_s": busy.value, "launches": la.value,
                "ticks": ti.value}

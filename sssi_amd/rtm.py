"""Reverse-time migration (2-D CPML) over the C ABI (include/sssi.h, sssi_rtm2d_*).

Mirrors the imaging loop of applications/rtm_2d_acoustic/mainRtmTimeCpmlFor2dAw.m:
per shot, dataTrue = fwd(V), dataSmooth = fwd(VS) with source-wavefield
checkpoints, back-propagation of dataTrue - dataSmooth in V with the
cross-correlation imaging condition, and an FP64 stack of M ./ (eps + s2).
"""
import ctypes

import numpy as np

from ._lib import _call, _dp, _f64, lib

CK_FORMATS = {"fp32": 1, "fp16": 2, "bf16": 3}


class Rtm2D:
    """A batch of RTM shots on sub-windows of one model.

    V, VS: physical models indexed [iz, ix]; each shot migrates the window
    [x_begin, x_begin + x_count) extended by nBoundary cells (left, right, below).
    """

    def __init__(self, V, VS, nDiffOrder, nBoundary, dz, dx, dt, nt, img_stride=1,
                 ck_format="fp32", eps=1e-3):
        V = _f64(V)
        VS = _f64(VS)
        if V.shape != VS.shape:
            raise ValueError("V and VS must have the same shape")
        self.nz, self.nx = V.shape
        self.nB = int(nBoundary)
        self.nt = int(nt)
        h = ctypes.c_void_p()
        _call("sssi_rtm2d_create", _dp(V), _dp(VS), self.nz, self.nx, int(nDiffOrder), self.nB,
              float(dz), float(dx), float(dt), self.nt, int(img_stride), CK_FORMATS[ck_format],
              float(eps), ctypes.byref(h))
        self._h = h
        self.windows = []

    def close(self):
        if self._h:
            lib().sssi_rtm2d_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_shot(self, x_begin, x_count, src_ix, wavelet):
        w = np.ascontiguousarray(wavelet, dtype=np.float64)
        if w.size != self.nt:
            raise ValueError("wavelet must have nt samples")
        sid = ctypes.c_int(-1)
        _call("sssi_rtm2d_add_shot", self._h, int(x_begin), int(x_count), int(src_ix), _dp(w),
              ctypes.byref(sid))
        self.windows.append((int(x_begin), int(x_count)))
        return sid.value

    def precompute_data(self, batched=True):
        _call("sssi_rtm2d_precompute_data", self._h, int(bool(batched)))

    def migrate(self, batched=True):
        """Run the imaging job of every shot; returns (milliseconds, launches)."""
        ms = ctypes.c_double(0)
        n = ctypes.c_int64(0)
        _call("sssi_rtm2d_migrate", self._h, int(bool(batched)), ctypes.byref(ms), ctypes.byref(n))
        return ms.value, n.value

    def stack(self):
        out = np.zeros((self.nz, self.nx), order="F")
        _call("sssi_rtm2d_get_stack", self._h, _dp(out))
        return out

    def shot(self, k):
        """Per-shot image, M, s2 on the extended window and dataTrue (nx_ext, nt)."""
        xb, xc = self.windows[k]
        ze, xe = self.nz + self.nB, xc + 2 * self.nB
        img, num, den = (np.zeros((ze, xe), order="F") for _ in range(3))
        dtrue = np.zeros((xe, self.nt), order="F")
        _call("sssi_rtm2d_get_shot", self._h, int(k), _dp(img), _dp(num), _dp(den), _dp(dtrue))
        return {"image": img, "num": num, "den": den, "data_true": dtrue}

    def set_cache(self, directory):
        """Durable per-shot results in `directory` (created if missing; None turns
        caching off): dataTrue<xs>.mat and image<xs>.mat are loaded when they
        exist and written when they do not, StackedRTM.mat after a migration
        (xs = x_begin + src_ix + 1, the shot's 1-based source column)."""
        _call("sssi_rtm2d_set_cache", self._h, None if directory is None else
              str(directory).encode())

    def cache_stats(self):
        """(dataTrue loaded, images loaded) in the last run, files written so far."""
        a, b, c = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
        _call("sssi_rtm2d_cache_stats", self._h, ctypes.byref(a), ctypes.byref(b),
              ctypes.byref(c))
        return {"loaded_data": a.value, "loaded_images": b.value, "saved": c.value}

    def device_bytes(self):
        b = ctypes.c_double(0)
        _call("sssi_rtm2d_device_bytes", self._h, ctypes.byref(b))
        return b.value

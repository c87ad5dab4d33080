"""ctypes binding of the run-loop driver (include/hc_model.h)."""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import hc


class hcm_options(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("fused", ctypes.c_int), ("quiet", ctypes.c_int),
                ("write_files", ctypes.c_int), ("flat_dz", ctypes.c_double)]


def _lib():
    L = hc.lib()
    if not getattr(L, "_hcm_ready", False):
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.hcm_create.argtypes = [ctypes.POINTER(vp), ctypes.c_char_p, ctypes.c_char_p,
                                 ctypes.POINTER(hcm_options)]
        L.hcm_destroy.argtypes = [vp]
        L.hcm_last_error.argtypes = [vp]
        L.hcm_last_error.restype = ctypes.c_char_p
        L.hcm_iterate.argtypes = [vp, i, vp, vp, vp, ctypes.POINTER(i), ctypes.POINTER(i)]
        L.hcm_download.argtypes = [vp, i, vp, ctypes.c_size_t]
        L.hcm_field_count.argtypes = [vp, i]
        L.hcm_field_count.restype = ctypes.c_size_t
        L.hcm_scalar.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(d)]
        L.hcm_series.argtypes = [vp, vp, i]
        L._hcm_ready = True
    return L


class Model:
    """The reference run loop on the device (src/main.cpp:103-204)."""

    def __init__(self, path: str, params_file: str, precision: int = hc.HC_FP64,
                 fused: bool = True, write_files: bool = False, flat_dz: float = 0.0):
        self.L = _lib()
        if not path.endswith("/"):
            path = path + "/"
        opt = hcm_options(precision, 1 if fused else 0, 1, 1 if write_files else 0, flat_dz)
        h = ctypes.c_void_p()
        rc = self.L.hcm_create(ctypes.byref(h), path.encode(), params_file.encode(), ctypes.byref(opt))
        self.h = h
        if rc != 0:
            msg = self.L.hcm_last_error(h).decode()
            self.L.hcm_destroy(h)
            self.h = ctypes.c_void_p()
            raise hc.HCError(rc, msg)
        self.imax = int(self.scalar("imax"))
        self.jmax = int(self.scalar("jmax"))

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.L.hcm_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def iterate(self, n: int):
        tt = np.zeros(n)
        cy = np.zeros(n)
        dt = np.zeros(n)
        done = ctypes.c_int()
        fin = ctypes.c_int()
        rc = self.L.hcm_iterate(self.h, n, tt.ctypes.data, cy.ctypes.data, dt.ctypes.data,
                                ctypes.byref(done), ctypes.byref(fin))
        if rc != 0:
            raise hc.HCError(rc, self.L.hcm_last_error(self.h).decode())
        k = done.value
        self.finished = bool(fin.value)
        return tt[:k], cy[:k], dt[:k]

    def scalar(self, name: str) -> float:
        v = ctypes.c_double()
        if self.L.hcm_scalar(self.h, name.encode(), ctypes.byref(v)) != 0:
            raise KeyError(name)
        return v.value

    def download(self, name: str) -> np.ndarray:
        fid = hc.FIELDS[name]
        k = hc.PLANES.get(name, 1)
        shape = (self.imax + 2, self.jmax + 2) + ((k,) if k > 1 else ())
        a = np.zeros(shape, dtype=np.int32 if name in hc.INT_FIELDS else np.float64)
        rc = self.L.hcm_download(self.h, fid, a.ctypes.data, a.size)
        if rc != 0:
            raise hc.HCError(rc, self.L.hcm_last_error(self.h).decode())
        return a

    def series(self) -> np.ndarray:
        rows = self.L.hcm_series(self.h, None, 0)
        out = np.zeros((rows, 14))
        if rows:
            self.L.hcm_series(self.h, out.ctypes.data, rows)
        return out

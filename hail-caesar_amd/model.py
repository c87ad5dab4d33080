"""ctypes binding of the run-loop driver (include/hc_model.h)."""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

import contextlib
import os

from . import hc


@contextlib.contextmanager
def _in_dir(path):
    """Relative read/write paths in a parameter file are relative to the
    working directory, as in the reference's CLI; run inside `path`."""
    if path is None:
        yield
        return
    old = os.getcwd()
    os.chdir(path)
    try:
        yield
    finally:
        os.chdir(old)


class hcm_options(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("fused", ctypes.c_int), ("quiet", ctypes.c_int),
                ("write_files", ctypes.c_int), ("flat_dz", ctypes.c_double),
                ("device_loop", ctypes.c_int), ("use_graph", ctypes.c_int),
                ("elev_fixed", ctypes.c_int), ("dt_floor", ctypes.c_double)]


class hcm_arrays(ctypes.Structure):
    _fields_ = [("nrows", ctypes.c_int), ("ncols", ctypes.c_int), ("DX", ctypes.c_double),
                ("xll", ctypes.c_double), ("yll", ctypes.c_double), ("dem", ctypes.c_void_p),
                ("mannings", ctypes.c_void_p), ("water_init", ctypes.c_void_p),
                ("hydroindex", ctypes.c_void_p), ("rain", ctypes.c_void_p),
                ("rain_rows", ctypes.c_int), ("rain_cols", ctypes.c_int),
                ("graindata", ctypes.c_void_p)]


def _lib():
    L = hc.lib()
    if not getattr(L, "_hcm_ready", False):
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.hcm_create.argtypes = [ctypes.POINTER(vp), ctypes.c_char_p, ctypes.c_char_p,
                                 ctypes.POINTER(hcm_options)]
        L.hcm_create_arrays.argtypes = [ctypes.POINTER(vp), ctypes.c_char_p,
                                        ctypes.POINTER(hcm_arrays), ctypes.POINTER(hcm_options)]
        L.hcm_destroy.argtypes = [vp]
        L.hcm_last_error.argtypes = [vp]
        L.hcm_last_error.restype = ctypes.c_char_p
        L.hcm_iterate.argtypes = [vp, i, vp, vp, vp, ctypes.POINTER(i), ctypes.POINTER(i)]
        L.hcm_download.argtypes = [vp, i, vp, ctypes.c_size_t]
        L.hcm_field_count.argtypes = [vp, i]
        L.hcm_field_count.restype = ctypes.c_size_t
        L.hcm_scalar.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(d)]
        L.hcm_series.argtypes = [vp, vp, i]
        L.hcm_set_graph.argtypes = [vp, i]
        L.hcm_context.argtypes = [vp]
        L.hcm_context.restype = vp
        L.hcm_loop_begin.argtypes = [vp]
        L.hcm_loop_collect.argtypes = [vp, i, vp, vp, vp, ctypes.POINTER(i), ctypes.POINTER(i)]
        L.hcm_checkpoint.argtypes = [vp, ctypes.c_char_p]
        L.hcm_restart.argtypes = [vp, ctypes.c_char_p]
        L._hcm_ready = True
    return L


class Model:
    """The reference run loop on the device (src/main.cpp:103-204)."""

    def __init__(self, path: Optional[str], params_file: Optional[str], precision: int = hc.HC_FP64,
                 fused: bool = True, write_files: bool = False, flat_dz: float = 0.0,
                 device_loop: bool = False, use_graph: bool = False, arrays: Optional[dict] = None,
                 params: Optional[dict] = None, elev_fixed: bool = False, dt_floor: float = 0.0):
        """Files (path + params_file, as the reference's CLI) or, with
        ``arrays`` and ``params``, in-memory inputs: arrays holds 'dem' and
        optionally 'mannings', 'water_init', 'hydroindex', 'rain' (rows x
        cells, mm/h), 'graindata' (nrows x ncols x 101: grain[0..10] and
        strata[10][9] per cell, the grain data file's values) plus 'DX'.

        dt_floor: a body whose flow timestep would fall below it (seconds)
        stops the run with HC_ERR_UNSTABLE, as does a depth gone non-finite
        or below -1e-4 m; 0 selects 1e-4 s, a negative value turns the
        timestep test off."""
        self.L = _lib()
        opt = hcm_options(precision, 1 if fused else 0, 1, 1 if write_files else 0, flat_dz,
                          1 if device_loop else 0, 1 if use_graph else 0, 1 if elev_fixed else 0,
                          dt_floor)
        h = ctypes.c_void_p()
        self.workdir = None
        if arrays is None:
            self.workdir = os.path.abspath(path)
            with _in_dir(self.workdir):
                rc = self.L.hcm_create(ctypes.byref(h), b"./", params_file.encode(), ctypes.byref(opt))
        else:
            keep = {}

            def ptr(name, dtype):
                if arrays.get(name) is None:
                    return None
                a = np.ascontiguousarray(arrays[name], dtype=dtype)
                keep[name] = a
                return a.ctypes.data
            dem = np.ascontiguousarray(arrays["dem"], dtype=np.float64)
            keep["dem"] = dem
            rain = arrays.get("rain")
            rr = rc_ = 0
            if rain is not None:
                rain = np.ascontiguousarray(np.atleast_2d(np.asarray(rain, dtype=np.float32).T).T)
                if rain.ndim == 1:
                    rain = rain[:, None]
                keep["rain"] = rain
                rr, rc_ = rain.shape
            ar = hcm_arrays(dem.shape[0], dem.shape[1], float(arrays["DX"]), 0.0, 0.0, dem.ctypes.data,
                            ptr("mannings", np.float64), ptr("water_init", np.float64),
                            ptr("hydroindex", np.int32), rain.ctypes.data if rain is not None else None,
                            rr, rc_, ptr("graindata", np.float64))
            text = "".join(f"{k}: {v}\n" for k, v in params.items())
            rc = self.L.hcm_create_arrays(ctypes.byref(h), text.encode(), ctypes.byref(ar), ctypes.byref(opt))
        self.h = h
        if rc != 0:
            msg = self.L.hcm_last_error(h).decode()
            self.L.hcm_destroy(h)
            self.h = ctypes.c_void_p()
            raise hc.HCError(rc, msg)
        self.imax = int(self.scalar("imax"))
        self.jmax = int(self.scalar("jmax"))
        hc.track(self)

    _close_rank = 2

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            if not hc._exiting:
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
        with _in_dir(self.workdir):
            rc = self.L.hcm_iterate(self.h, n, tt.ctypes.data, cy.ctypes.data, dt.ctypes.data,
                                    ctypes.byref(done), ctypes.byref(fin))
        k = done.value
        if rc != 0:  # a stopped run keeps the records of the bodies that ran
            err = hc.HCError(rc, self.L.hcm_last_error(self.h).decode())
            err.results = (tt[:k], cy[:k], dt[:k])
            raise err
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

    def checkpoint(self, path: str) -> None:
        """Exact binary checkpoint of a host-stepped run (hcm_checkpoint)."""
        rc = self.L.hcm_checkpoint(self.h, os.fsencode(path))
        if rc != 0:
            raise hc.HCError(rc, self.L.hcm_last_error(self.h).decode())

    def restart(self, path: str) -> None:
        """Continue from a checkpoint written by a model of the same inputs
        (hcm_restart); a failed restart leaves the model as it was."""
        rc = self.L.hcm_restart(self.h, os.fsencode(path))
        if rc != 0:
            raise hc.HCError(rc, self.L.hcm_last_error(self.h).decode())

    def set_graph(self, on: bool) -> None:
        self.L.hcm_set_graph(self.h, 1 if on else 0)

    def context(self) -> "hc.Context":
        return hc.Context.borrow(self.L.hcm_context(self.h), self.imax, self.jmax)

    def synchronize(self) -> None:
        self.context().synchronize()


class Ensemble:
    """Independent catchments stepped together by the batched kernels
    (include/hc.h, hc_ens_*).  Every member is a device-loop Model; after each
    ensemble run every member's clock, records and output series are brought
    up to date as if it had been stepped alone."""

    def __init__(self, models):
        self.L = _lib()
        self.models = list(models)
        for m in self.models:
            if self.L.hcm_loop_begin(m.h) != 0:
                raise hc.HCError(hc.HC_ERR_STATE, self.L.hcm_last_error(m.h).decode())
        ctxs = (ctypes.c_void_p * len(self.models))(*[self.L.hcm_context(m.h) for m in self.models])
        e = ctypes.c_void_p()
        rc = self.L.hc_ens_create(ctypes.byref(e), ctxs, len(self.models))
        if rc != 0:
            raise hc.HCError(rc, "hc_ens_create failed")
        self.e = e
        hc.track(self)

    _close_rank = 0

    def close(self):
        if getattr(self, "e", None) and self.e.value:
            if not hc._exiting:
                self.L.hc_ens_destroy(self.e)
            self.e = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise hc.HCError(rc, self.L.hc_ens_last_error(self.e).decode())

    def run(self, steps: int, use_graph: bool = True):
        """Enqueue `steps` ensemble steps and wait (no per-member collection)."""
        self._check(self.L.hc_ens_run(self.e, steps, 1 if use_graph else 0))
        self._check(self.L.hc_ens_sync(self.e))

    def collect(self, steps: int):
        """Per member (temptot, cycle, dt_flow) of the last run's steps."""
        out = []
        first = None  # the first member that stopped on a failure
        for i, m in enumerate(self.models):
            tt, cy, dt = np.zeros(steps), np.zeros(steps), np.zeros(steps)
            done, fin = ctypes.c_int(), ctypes.c_int()
            with _in_dir(m.workdir):
                rc = self.L.hcm_loop_collect(m.h, steps, tt.ctypes.data, cy.ctypes.data,
                                             dt.ctypes.data, ctypes.byref(done), ctypes.byref(fin))
            k = done.value
            if rc != 0:
                if rc not in (hc.HC_ERR_UNSTABLE, hc.HC_ERR_HALTED):
                    raise hc.HCError(rc, self.L.hcm_last_error(m.h).decode())
                if first is None:
                    first = hc.HCError(rc, f"member {i}: " + self.L.hcm_last_error(m.h).decode())
                    first.member = i
            else:
                m.finished = bool(fin.value)
            out.append((tt[:k], cy[:k], dt[:k]))
        if first is not None:  # every member's records are collected first
            first.results = out
            raise first
        return out

    def iterate(self, steps: int, use_graph: bool = True):
        self.run(steps, use_graph)
        return self.collect(steps)

    def tune(self):
        r, b, us = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        self._check(self.L.hc_ens_tune(self.e, ctypes.byref(r), ctypes.byref(b), ctypes.byref(us)))
        return r.value, b.value, us.value

    def set_tiles(self, rows: int, blocks: int):
        self._check(self.L.hc_ens_set_tiles(self.e, rows, blocks))

    def bench_fused(self, reps: int = 10) -> float:
        us = ctypes.c_double()
        self._check(self.L.hc_ens_bench_fused(self.e, reps, ctypes.byref(us)))
        return us.value

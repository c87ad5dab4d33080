"""ctypes binding of the C-ABI in include/hc.h (libhc.so).

Host arrays use the reference's padded layout ((imax+2) x (jmax+2), x=row,
y=column contiguous; multi-plane fields planes-last).  The library must have
been built (``python hail-caesar_amd/build.py``); importing this module on a
machine without it raises, and nothing here falls back to a CPU path.
"""
from __future__ import annotations

import atexit
import ctypes
import os
import weakref
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhc.so")

HC_OK, HC_ERR_ARG, HC_ERR_HIP, HC_ERR_STATE, HC_ERR_UNSTABLE, HC_ERR_HALTED = 0, -1, -2, -3, -4, -5
HC_FP64, HC_FP32 = 0, 1
HC_WILCOCK, HC_EINSTEIN = 1, 2
HC_NOALLOC_MAX_POINTS = 16384

FIELDS = {
    "water_depth": 0, "elev": 1, "qx": 2, "qy": 3, "mannings": 4, "Vsusptot": 5,
    "qxs": 6, "qys": 7, "vel_dir": 8, "mask": 9, "grain": 10, "strata": 11,
    "hasgrain": 12, "bedrock": 13, "Tau": 14, "Vel": 15, "ss": 16, "sr": 17,
    "sl": 18, "su": 19, "sd": 20,
}
PLANES = {"vel_dir": 9, "grain": 11, "strata": 90, "sr": 10, "sl": 10, "su": 10, "sd": 10}
INT_FIELDS = {"mask", "hasgrain"}

# exported symbols (must match include/hc.h)
SYMBOLS = ["hc_create", "hc_destroy", "hc_last_error", "hc_field_count", "hc_upload",
           "hc_download", "hc_flow_route", "hc_depth_update", "hc_flow_depth",
           "hc_scan_area", "hc_water_flux_out", "hc_add_point_inputs",
           "hc_add_point_inputs_noalloc", "hc_points_set", "hc_points_add", "hc_erode",
           "hc_global_landsliding", "hc_creep", "hc_in_channel_landsliding",
           "hc_grain_index_set", "hc_grain_index_get",
           "hc_reach_inputs", "hc_runoff_set", "hc_runoff_add", "hc_synchronize", "hc_check_health", "hc_state_bytes", "hc_state_save",
           "hc_state_load", "hc_loop_begin", "hc_loop_run",
           "hc_loop_sync", "hc_loop_profile", "hc_loop_kernel_time", "hc_tune_tiles",
           "hc_set_tiles", "hc_bench_fused", "hc_ens_create", "hc_ens_destroy",
           "hc_ens_last_error", "hc_ens_run", "hc_ens_sync", "hc_ens_tune", "hc_ens_set_tiles",
           "hc_ens_bench_fused", "hc_adj_create", "hc_adj_destroy", "hc_adj_last_error",
           "hc_adj_set_target", "hc_adj_forward", "hc_adj_backward"]
MODEL_SYMBOLS = ["hcm_create", "hcm_create_arrays", "hcm_destroy", "hcm_last_error",
                 "hcm_iterate", "hcm_download", "hcm_field_count", "hcm_scalar",
                 "hcm_series", "hcm_set_graph", "hcm_context", "hcm_loop_begin",
                 "hcm_loop_collect", "hcm_checkpoint", "hcm_restart"]
RF_MAX = 17


class hc_grid(ctypes.Structure):
    _fields_ = [("imax", ctypes.c_int), ("jmax", ctypes.c_int), ("DX", ctypes.c_double),
                ("precision", ctypes.c_int), ("sediment", ctypes.c_int)]


class hc_params(ctypes.Structure):
    _fields_ = [
        ("courant_number", ctypes.c_double), ("froude_limit", ctypes.c_double),
        ("hflow_threshold", ctypes.c_double), ("edgeslope", ctypes.c_double),
        ("mannings", ctypes.c_double), ("spatially_var_mannings", ctypes.c_int),
        ("water_depth_erosion_threshold", ctypes.c_double), ("ERODEFACTOR", ctypes.c_double),
        ("active", ctypes.c_double), ("Csuspmax", ctypes.c_double), ("max_vel", ctypes.c_double),
        ("transport_law", ctypes.c_int), ("d", ctypes.c_double * 10),
        ("dprop", ctypes.c_double * 11), ("fallVelocity", ctypes.c_double * 11),
        ("bedrock_erosion_threshold", ctypes.c_double), ("bedrock_erosion_rate", ctypes.c_double),
        ("lateral_constant", ctypes.c_double), ("chann_lateral_erosion", ctypes.c_double),
        ("veg_lat_restriction", ctypes.c_double), ("flat_dz", ctypes.c_double),
        ("elev_fixed", ctypes.c_int),
    ]


class hc_loop_init(ctypes.Structure):
    _fields_ = [
        ("time_step", ctypes.c_double), ("cycle", ctypes.c_double), ("maxdepth", ctypes.c_double),
        ("waterinput", ctypes.c_double), ("waterOut", ctypes.c_double), ("time_1", ctypes.c_double),
        ("baseflow", ctypes.c_double), ("counter", ctypes.c_int),
        ("courant_number", ctypes.c_double), ("in_out_difference_allowed", ctypes.c_double),
        ("output_file_save_interval", ctypes.c_double), ("ERODEFACTOR", ctypes.c_double),
        ("MIN_Q", ctypes.c_double), ("MIN_Q_MAXVAL", ctypes.c_double),
        ("max_time_step", ctypes.c_int), ("maxcycle", ctypes.c_int), ("rfnum", ctypes.c_int),
        ("j_mean", ctypes.c_double * RF_MAX), ("old_j_mean", ctypes.c_double * RF_MAX),
        ("new_j_mean", ctypes.c_double * RF_MAX), ("nActual", ctypes.c_int * RF_MAX),
        ("input_counter", ctypes.c_int * RF_MAX), ("topmodel", ctypes.c_void_p),
        ("ntop", ctypes.c_int), ("area", ctypes.c_void_p), ("rfarea", ctypes.c_void_p),
        ("in_x", ctypes.c_void_p), ("in_y", ctypes.c_void_p), ("npoints", ctypes.c_int),
        ("first_x", ctypes.c_int), ("first_y", ctypes.c_int), ("dt_floor", ctypes.c_double),
    ]


class hc_loop_record(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("cycle", "previous", "old_cycle", "new_cycle",
                                                 "temptot", "dt_flow", "time_step", "waterinput")]
    _fields_ += [("j_mean", ctypes.c_double * RF_MAX)]


class hc_loop_state(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("time_step", "cycle", "maxdepth", "temptot",
                                                 "waterinput", "waterOut", "time_1", "baseflow")]
    _fields_ += [("counter", ctypes.c_int), ("finished", ctypes.c_int), ("npoints", ctypes.c_int),
                 ("error", ctypes.c_int), ("j_mean", ctypes.c_double * RF_MAX),
                 ("old_j_mean", ctypes.c_double * RF_MAX), ("new_j_mean", ctypes.c_double * RF_MAX)]


RECORD_DTYPE = np.dtype([(n, np.float64) for n in ("cycle", "previous", "old_cycle", "new_cycle",
                                                     "temptot", "dt_flow", "time_step", "waterinput")]
                        + [("j_mean", np.float64, (RF_MAX,))])


class HCError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libhc error {code}: {msg}")
        self.code = code


_lib = None
# Every live handle object (contexts, adjoints, models, ensembles), closed by
# _close_all at interpreter exit while the HIP runtime is still whole.
_live = weakref.WeakSet()
_exiting = False


def track(obj) -> None:
    """Register an object with a close() that frees device resources; lower
    `_close_rank` is closed first (borrowers before what they borrow)."""
    _live.add(obj)


def _close_all() -> None:
    """atexit: close every live handle, borrowers first (ensembles borrow
    member contexts, adjoints borrow a context).  Registered after torch's
    own exit handlers (its import comes first), so it runs before them; a
    handle left open into the runtime's static teardown aborts the process
    (std::bad_variant_access).  Later finalizers find their handles closed."""
    global _exiting
    for obj in sorted(list(_live), key=lambda o: getattr(o, "_close_rank", 9)):
        try:
            obj.close()
        except Exception:
            pass
    _exiting = True


def hip_runtimes_mapped() -> set:
    """Distinct HIP runtime files (libamdhip64) mapped into this process."""
    paths = set()
    try:
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64" in line:
                    paths.add(os.path.realpath(line.split()[-1]))
    except OSError:
        pass
    return paths


def _choose_runtime() -> None:
    """Decide the HIP runtime libhc.so binds to, before it is loaded.

    The torch wheel bundles its own HIP runtime (torch/lib/libamdhip64.so,
    soname libamdhip64.so.7) and loads it by a name that does not match the
    system runtime's file.  Loaded after libhc.so has initialised the system
    runtime (/opt/rocm/lib), torch maps a second copy, and the two copies
    tear each other down at exit (std::bad_variant_access, abort).  Loaded
    first, its runtime satisfies libhc.so's libamdhip64.so.7 dependency and
    the process has exactly one.  So wherever torch is importable it is
    imported here, first; every Python caller (tests, benchmarks, the
    autograd harness) then runs libhc.so on the same runtime."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with hail-caesar_amd/build.py")
        _choose_runtime()
        L = ctypes.CDLL(LIB_PATH)
        rt = hip_runtimes_mapped()
        if len(rt) > 1:
            raise HCError(HC_ERR_STATE, "two HIP runtimes are mapped into this process "
                          f"({', '.join(sorted(rt))}); import torch before anything loads "
                          "the system HIP runtime")
        atexit.register(_close_all)
        vp, i, d, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_size_t
        L.hc_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(hc_grid), ctypes.POINTER(hc_params)]
        L.hc_destroy.argtypes = [vp]
        L.hc_last_error.argtypes = [vp]
        L.hc_last_error.restype = ctypes.c_char_p
        L.hc_field_count.argtypes = [vp, i]
        L.hc_field_count.restype = sz
        L.hc_upload.argtypes = [vp, i, vp, sz]
        L.hc_download.argtypes = [vp, i, vp, sz]
        L.hc_flow_route.argtypes = [vp, d]
        L.hc_depth_update.argtypes = [vp, d, ctypes.POINTER(d)]
        L.hc_flow_depth.argtypes = [vp, d, ctypes.POINTER(d)]
        L.hc_scan_area.argtypes = [vp]
        L.hc_water_flux_out.argtypes = [vp, d, ctypes.POINTER(d)]
        L.hc_add_point_inputs.argtypes = [vp, vp, vp, vp, i]
        L.hc_add_point_inputs_noalloc.argtypes = [vp, vp, vp, vp, i]
        L.hc_points_set.argtypes = [vp, vp, vp, vp, i]
        L.hc_points_add.argtypes = [vp, vp, i]
        L.hc_erode.argtypes = [vp, d, ctypes.POINTER(d), ctypes.POINTER(d), vp]
        L.hc_global_landsliding.argtypes = [vp, d, ctypes.POINTER(i)]
        L.hc_creep.argtypes = [vp, d, d]
        L.hc_in_channel_landsliding.argtypes = [vp, d]
        L.hc_grain_index_set.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_longlong]
        L.hc_grain_index_get.argtypes = [vp, vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_longlong)]
        L.hc_synchronize.argtypes = [vp]
        L.hc_check_health.argtypes = [vp, ctypes.POINTER(ctypes.c_longlong)]
        L.hc_state_bytes.argtypes = [vp, ctypes.POINTER(ctypes.c_size_t)]
        L.hc_runoff_set.argtypes = [vp, vp, vp]
        L.hc_reach_inputs.argtypes = [vp, i, vp, vp, vp, vp, d]
        L.hc_runoff_add.argtypes = [vp, d, d, d]
        L.hc_state_save.argtypes = [vp, vp, ctypes.c_size_t]
        L.hc_state_load.argtypes = [vp, vp, ctypes.c_size_t]
        L.hc_loop_begin.argtypes = [vp, ctypes.POINTER(hc_loop_init)]
        L.hc_loop_run.argtypes = [vp, i, i]
        L.hc_loop_sync.argtypes = [vp, ctypes.POINTER(hc_loop_state), vp, i, ctypes.POINTER(i)]
        L.hc_tune_tiles.argtypes = [vp, d, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(d)]
        L.hc_set_tiles.argtypes = [vp, i, i]
        L.hc_bench_fused.argtypes = [vp, d, i, ctypes.POINTER(d)]
        L.hc_loop_profile.argtypes = [vp, i]
        L.hc_loop_kernel_time.argtypes = [vp, ctypes.POINTER(d), ctypes.POINTER(i)]
        L.hc_ens_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(vp), i]
        L.hc_ens_destroy.argtypes = [vp]
        L.hc_ens_last_error.argtypes = [vp]
        L.hc_ens_last_error.restype = ctypes.c_char_p
        L.hc_ens_run.argtypes = [vp, i, i]
        L.hc_ens_sync.argtypes = [vp]
        L.hc_ens_tune.argtypes = [vp, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(d)]
        L.hc_ens_set_tiles.argtypes = [vp, i, i]
        L.hc_ens_bench_fused.argtypes = [vp, i, ctypes.POINTER(d)]
        L.hc_adj_create.argtypes = [ctypes.POINTER(vp), vp, i, d, i, vp, i, vp]
        L.hc_adj_destroy.argtypes = [vp]
        L.hc_adj_last_error.argtypes = [vp]
        L.hc_adj_last_error.restype = ctypes.c_char_p
        L.hc_adj_set_target.argtypes = [vp, vp]
        L.hc_adj_forward.argtypes = [vp, vp, ctypes.POINTER(d), vp]
        L.hc_adj_backward.argtypes = [vp, vp, vp]
        _lib = L
    return _lib


def default_params(**kw) -> hc_params:
    """Reference defaults (LSDCatchmentModel.hpp:437-568) overridden by kw."""
    p = hc_params()
    p.courant_number = 0.7
    p.froude_limit = 0.8
    p.hflow_threshold = 0.00001
    p.edgeslope = 0.001
    p.mannings = 0.04
    p.spatially_var_mannings = 0
    p.water_depth_erosion_threshold = 0.01
    p.ERODEFACTOR = 0.05
    p.active = 0.2
    p.Csuspmax = 0.05
    p.max_vel = 5.0
    p.transport_law = HC_WILCOCK
    d = [0.0, 0.000065, 0.001, 0.002, 0.004, 0.008, 0.016, 0.032, 0.064, 0.128]
    dprop = [0.0, 0.05, 0.05, 0.15, 0.225, 0.25, 0.1, 0.075, 0.05, 0.05, 0.0]
    fall = [0.0, 0.033, 0.109, 0.164, 0.237, 0.338, 0.479, 0.678, 0.959, 1.357, 0.0]
    for k in range(10):
        p.d[k] = d[k]
    for k in range(11):
        p.dprop[k] = dprop[k]
        p.fallVelocity[k] = fall[k]
    p.bedrock_erosion_threshold = 0.0
    p.bedrock_erosion_rate = 0.0
    p.lateral_constant = 0.0000002
    p.chann_lateral_erosion = 20.0
    p.veg_lat_restriction = 0.1
    p.flat_dz = 0.0
    p.elev_fixed = 0
    for k, v in kw.items():
        if k in ("d", "dprop", "fallVelocity"):
            arr = getattr(p, k)
            for j, x in enumerate(v):
                arr[j] = x
        else:
            setattr(p, k, v)
    return p


class Context:
    """One device context (hc_ctx*)."""

    @classmethod
    def borrow(cls, handle: int, imax: int, jmax: int):
        """Wrap a context owned by someone else (e.g. a driver model)."""
        c = cls.__new__(cls)
        c.L = lib()
        c.h = ctypes.c_void_p(handle)
        c.imax, c.jmax = imax, jmax
        c.shape = (imax + 2, jmax + 2)
        c._borrowed = True
        return c

    def __init__(self, imax: int, jmax: int, DX: float, params: hc_params,
                 precision: int = HC_FP64, sediment: bool = False):
        self._borrowed = False
        self.L = lib()
        g = hc_grid(imax, jmax, DX, precision, 1 if sediment else 0)
        h = ctypes.c_void_p()
        rc = self.L.hc_create(ctypes.byref(h), ctypes.byref(g), ctypes.byref(params))
        if rc != HC_OK:
            raise HCError(rc, "hc_create failed")
        self.h = h
        self.imax, self.jmax, self.DX = imax, jmax, DX
        self.shape = (imax + 2, jmax + 2)
        self.precision = precision
        self.sediment = sediment
        track(self)

    _close_rank = 3

    def close(self):
        if getattr(self, "_borrowed", False):
            return
        if getattr(self, "h", None) and self.h.value:
            if not _exiting:
                self.L.hc_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc != HC_OK:
            raise HCError(rc, self.L.hc_last_error(self.h).decode())

    def upload(self, name: str, arr: np.ndarray) -> None:
        fid = FIELDS[name]
        if name in INT_FIELDS:
            a = np.ascontiguousarray(arr, dtype=np.int32)
        else:
            a = np.ascontiguousarray(arr, dtype=np.float64)
        self._check(self.L.hc_upload(self.h, fid, a.ctypes.data, a.size))

    def download(self, name: str) -> np.ndarray:
        fid = FIELDS[name]
        k = PLANES.get(name, 1)
        shape = self.shape if k == 1 else self.shape + (k,)
        a = np.zeros(shape, dtype=np.int32 if name in INT_FIELDS else np.float64)
        self._check(self.L.hc_download(self.h, fid, a.ctypes.data, a.size))
        return a

    def flow_route(self, dt: float) -> None:
        self._check(self.L.hc_flow_route(self.h, dt))

    def depth_update(self, dt: float) -> float:
        md = ctypes.c_double()
        self._check(self.L.hc_depth_update(self.h, dt, ctypes.byref(md)))
        return md.value

    def flow_depth(self, dt: float) -> float:
        md = ctypes.c_double()
        self._check(self.L.hc_flow_depth(self.h, dt, ctypes.byref(md)))
        return md.value

    def scan_area(self) -> None:
        self._check(self.L.hc_scan_area(self.h))

    def water_flux_out(self, time_step: float) -> float:
        t = ctypes.c_double()
        self._check(self.L.hc_water_flux_out(self.h, time_step, ctypes.byref(t)))
        return t.value

    def add_point_inputs(self, xs: Sequence[int], ys: Sequence[int], amts: Sequence[float]) -> None:
        x = np.ascontiguousarray(xs, dtype=np.int32)
        y = np.ascontiguousarray(ys, dtype=np.int32)
        a = np.ascontiguousarray(amts, dtype=np.float64)
        self._check(self.L.hc_add_point_inputs(self.h, x.ctypes.data, y.ctypes.data, a.ctypes.data, len(x)))

    def add_point_inputs_noalloc(self, xs: Sequence[int], ys: Sequence[int],
                                 amts: Sequence[float]) -> None:
        """hc_add_point_inputs_noalloc: the same adds, staged in the device's
        static array (at most HC_NOALLOC_MAX_POINTS points)."""
        x = np.ascontiguousarray(xs, dtype=np.int32)
        y = np.ascontiguousarray(ys, dtype=np.int32)
        a = np.ascontiguousarray(amts, dtype=np.float64)
        self._check(self.L.hc_add_point_inputs_noalloc(self.h, x.ctypes.data, y.ctypes.data,
                                                        a.ctypes.data, len(x)))

    def points_set(self, xs: Sequence[int], ys: Sequence[int], rain_cell: Sequence[int]) -> None:
        """hc_points_set: the persistent input-point list and each point's rain cell."""
        x = np.ascontiguousarray(xs, dtype=np.int32)
        y = np.ascontiguousarray(ys, dtype=np.int32)
        r = np.ascontiguousarray(rain_cell, dtype=np.int32)
        self._check(self.L.hc_points_set(self.h, x.ctypes.data, y.ctypes.data, r.ctypes.data, len(x)))

    def points_add(self, amt_by_rain_cell: Sequence[float]) -> None:
        """hc_points_add: every listed point takes its rain cell's amount."""
        a = np.ascontiguousarray(amt_by_rain_cell, dtype=np.float64)
        self._check(self.L.hc_points_add(self.h, a.ctypes.data, len(a)))

    def grain_index_set(self, index: np.ndarray, rows_total: int) -> None:
        """hc_grain_index_set: the reference's row numbers of the initial
        grain rows ((imax+2) x (jmax+2), -9999 = none) and grain_array_tot."""
        a = np.ascontiguousarray(index, dtype=np.int32)
        assert a.shape == (self.imax + 2, self.jmax + 2)
        self._check(self.L.hc_grain_index_set(self.h, a.ctypes.data, a.size, int(rows_total)))

    def grain_index_get(self):
        """hc_grain_index_get: (index, grain_array_tot), the reference's row
        numbers of every cell's grain row."""
        a = np.zeros((self.imax + 2, self.jmax + 2), dtype=np.int32)
        tot = ctypes.c_longlong()
        self._check(self.L.hc_grain_index_get(self.h, a.ctypes.data, a.size, ctypes.byref(tot)))
        return a, tot.value

    def erode(self, mult: float, time_step: float):
        ts = ctypes.c_double(time_step)
        tb = ctypes.c_double()
        g = np.zeros(10, dtype=np.float32)
        self._check(self.L.hc_erode(self.h, mult, ctypes.byref(ts), ctypes.byref(tb), g.ctypes.data))
        return ts.value, tb.value, g

    def global_landsliding(self, failureangle: float) -> int:
        """LSDCatchmentModel::global_landsliding (LCM.cpp:5529-5654); returns the
        number of sweeps."""
        n = ctypes.c_int()
        self._check(self.L.hc_global_landsliding(self.h, failureangle, ctypes.byref(n)))
        return n.value

    def creep(self, time: float, creep_rate: float) -> None:
        """LSDCatchmentModel::creep(time) (LCM.cpp:5293-5427)."""
        self._check(self.L.hc_creep(self.h, time, creep_rate))

    def in_channel_landsliding(self, failureangle: float) -> None:
        """LSDCatchmentModel::in_channel_landsliding (LCM.cpp:5429-5528)."""
        self._check(self.L.hc_in_channel_landsliding(self.h, failureangle))

    def synchronize(self) -> None:
        self._check(self.L.hc_synchronize(self.h))

    def check_health(self) -> int:
        n = ctypes.c_longlong()
        rc = self.L.hc_check_health(self.h, ctypes.byref(n))
        if rc not in (HC_OK, HC_ERR_UNSTABLE):
            self._check(rc)
        return n.value

    # ---- spatially complex rainfall
    def runoff_set(self, old_j_mean: np.ndarray, new_j_mean: np.ndarray) -> None:
        """Upload the runoff grid's planes (hc_runoff_set)."""
        shape = (self.imax + 2, self.jmax + 2)
        a = np.ascontiguousarray(old_j_mean, dtype=np.float64)
        b = np.ascontiguousarray(new_j_mean, dtype=np.float64)
        if a.shape != shape or b.shape != shape:
            raise ValueError("runoff planes must be (imax+2, jmax+2)")
        self._check(self.L.hc_runoff_set(self.h, a.ctypes.data, b.ctypes.data))

    def runoff_add(self, time: float, flow_timestep: float, cap: float) -> None:
        """Add every cell's runoff over the step (hc_runoff_add)."""
        self._check(self.L.hc_runoff_add(self.h, time, flow_timestep, cap))

    # ---- exact state image
    def state_save(self) -> bytes:
        """The context's device state image (hc_state_save)."""
        n = ctypes.c_size_t()
        self._check(self.L.hc_state_bytes(self.h, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        self._check(self.L.hc_state_save(self.h, buf, n.value))
        return buf.raw

    def state_load(self, image: bytes) -> None:
        """Load an image written by state_save on a context of the same
        grid, precision and configuration (hc_state_load)."""
        buf = ctypes.create_string_buffer(bytes(image), len(image))
        self._check(self.L.hc_state_load(self.h, buf, len(image)))

    # ---- device-resident stepping
    def loop_begin(self, init: hc_loop_init) -> None:
        self._check(self.L.hc_loop_begin(self.h, ctypes.byref(init)))

    def loop_run(self, n: int, use_graph: bool = True) -> None:
        self._check(self.L.hc_loop_run(self.h, n, 1 if use_graph else 0))

    def loop_sync(self, max_records: int = 0):
        st = hc_loop_state()
        rec = np.zeros(max(1, max_records), dtype=RECORD_DTYPE)
        n = ctypes.c_int()
        rc = self.L.hc_loop_sync(self.h, ctypes.byref(st), rec.ctypes.data, max_records, ctypes.byref(n))
        rec = rec[:min(n.value, max_records)]
        if rc != HC_OK:  # a halted loop still reports its state and records
            err = HCError(rc, self.L.hc_last_error(self.h).decode())
            err.state, err.records = st, rec
            raise err
        return st, rec

    def loop_profile(self, on: bool) -> None:
        self._check(self.L.hc_loop_profile(self.h, 1 if on else 0))

    def loop_kernel_time(self):
        t = ctypes.c_double()
        n = ctypes.c_int()
        self._check(self.L.hc_loop_kernel_time(self.h, ctypes.byref(t), ctypes.byref(n)))
        return t.value, n.value

    def tune_tiles(self, dt: float):
        """(rows, strip, us) of the fastest fused-kernel tiling, now in use."""
        r = ctypes.c_int()
        st = ctypes.c_int()
        us = ctypes.c_double()
        self._check(self.L.hc_tune_tiles(self.h, dt, ctypes.byref(r), ctypes.byref(st), ctypes.byref(us)))
        return r.value, st.value, us.value

    def set_tiles(self, rows: int, strip: int = 1) -> None:
        self._check(self.L.hc_set_tiles(self.h, rows, strip))

    def bench_fused(self, dt: float, reps: int = 20) -> float:
        us = ctypes.c_double()
        self._check(self.L.hc_bench_fused(self.h, dt, reps, ctypes.byref(us)))
        return us.value


class Adjoint:
    """Outlet-hydrograph misfit and its gradients (include/hc.h, hc_adj_*) on
    an FP64 hydro-only context with spatially variable n.  The context's state
    when this object is made is the initial state of every forward run."""

    def __init__(self, ctx: "Context", nsteps: int, dt: float, block: np.ndarray,
                 rain: np.ndarray, checkpoint_every: int = 100):
        self.L = lib()
        self.ctx = ctx
        self.nsteps = nsteps
        self.block = np.ascontiguousarray(block, dtype=np.int32)
        assert self.block.shape == ctx.shape
        self.nb = int(self.block.max()) + 1
        rain = np.ascontiguousarray(rain, dtype=np.float64)
        assert rain.shape == (nsteps,)
        h = ctypes.c_void_p()
        rc = self.L.hc_adj_create(ctypes.byref(h), ctx.h, nsteps, dt, checkpoint_every,
                                  self.block.ctypes.data, self.nb, rain.ctypes.data)
        if rc != HC_OK:
            raise HCError(rc, "hc_adj_create failed")
        self.h = h
        track(self)

    _close_rank = 1

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            if not _exiting:
                self.L.hc_adj_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != HC_OK:
            raise HCError(rc, self.L.hc_adj_last_error(self.h).decode())

    def set_target(self, target: np.ndarray) -> None:
        t = np.ascontiguousarray(target, dtype=np.float64)
        assert t.shape == (self.nsteps,)
        self._check(self.L.hc_adj_set_target(self.h, t.ctypes.data))

    def forward(self, mult: Sequence[float]):
        m = np.ascontiguousarray(mult, dtype=np.float64)
        assert m.shape == (self.nb,)
        loss = ctypes.c_double()
        q = np.zeros(self.nsteps)
        self._check(self.L.hc_adj_forward(self.h, m.ctypes.data, ctypes.byref(loss), q.ctypes.data))
        return loss.value, q

    def backward(self):
        gn = np.zeros(self.ctx.shape)
        gm = np.zeros(self.nb)
        self._check(self.L.hc_adj_backward(self.h, gn.ctypes.data, gm.ctypes.data))
        return gn, gm

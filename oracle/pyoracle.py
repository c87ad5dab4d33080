"""Test infrastructure only: ctypes bridges to the oracle.

``OracleModel`` drives liboracle.so (the FP64 CPU restatement in
hc_oracle.cpp).  ``RefModel`` drives _ref/libhcref.so, the reference model
compiled from its own sources (only present where the reference tree was
available at build time).  Only tests/, the smoke check and the benchmark's
CPU baseline import this module.
"""
from __future__ import annotations

import ctypes
import os
from typing import Dict, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# HC_ORACLE_SO selects another build of the same source (the sanitizer
# builds of oracle/Makefile, driven by oracle/sanitize.py)
ORACLE_SO = os.environ.get("HC_ORACLE_SO") or os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libhcref.so")

# Reference defaults (LSDCatchmentModel.hpp:437-568) used when a key is absent.
DPROP = [0.0, 0.05, 0.05, 0.15, 0.225, 0.25, 0.1, 0.075, 0.05, 0.05, 0.0]
FALL_VELOCITY = [0.0, 0.033, 0.109, 0.164, 0.237, 0.338, 0.479, 0.678, 0.959, 1.357, 0.0]
REF_D = [0.0, 0.000065, 0.001, 0.002, 0.004, 0.008, 0.016, 0.032, 0.064, 0.128]


class orc_params(ctypes.Structure):
    _fields_ = [
        ("imax", ctypes.c_int), ("jmax", ctypes.c_int),
        ("DX", ctypes.c_double), ("courant_number", ctypes.c_double),
        ("froude_limit", ctypes.c_double), ("hflow_threshold", ctypes.c_double),
        ("edgeslope", ctypes.c_double), ("mannings", ctypes.c_double),
        ("spatially_var_mannings", ctypes.c_int),
        ("water_depth_erosion_threshold", ctypes.c_double),
        ("ERODEFACTOR", ctypes.c_double), ("active", ctypes.c_double),
        ("Csuspmax", ctypes.c_double), ("max_vel", ctypes.c_double),
        ("wilcock", ctypes.c_int), ("einstein", ctypes.c_int),
        ("d", ctypes.c_double * 10), ("dprop", ctypes.c_double * 11),
        ("fallVelocity", ctypes.c_double * 11),
        ("rfnum", ctypes.c_int), ("M", ctypes.c_double),
        ("rain_data_time_step", ctypes.c_double), ("max_time_step", ctypes.c_int),
        ("in_out_difference_allowed", ctypes.c_double),
        ("MIN_Q", ctypes.c_double), ("MIN_Q_MAXVAL", ctypes.c_double),
        ("maxcycle", ctypes.c_int), ("output_file_save_interval", ctypes.c_double),
        ("hydro_only", ctypes.c_int),
        ("bedrock_erosion_threshold", ctypes.c_double),
        ("bedrock_erosion_rate", ctypes.c_double),
        ("lateral_constant", ctypes.c_double),
        ("chann_lateral_erosion", ctypes.c_double),
        ("veg_lat_restriction", ctypes.c_double), ("vegTauCrit", ctypes.c_double),
        ("scan_interval", ctypes.c_int), ("nthreads", ctypes.c_int),
        ("spatially_complex_rainfall", ctypes.c_int), ("no_data_value", ctypes.c_double),
        ("reach_mode", ctypes.c_int), ("div_inputs", ctypes.c_int),
        ("reach_input_data_timestep", ctypes.c_int),
        ("creep_rate", ctypes.c_double), ("failureangle", ctypes.c_double),
    ]


def _atoi(s: str) -> int:
    s = s.strip()
    out = ""
    for i, ch in enumerate(s):
        if ch.isdigit() or (i == 0 and ch in "+-"):
            out += ch
        else:
            break
    try:
        return int(out)
    except ValueError:
        return 0


def _atof(s: str) -> float:
    try:
        return float(s.strip().split()[0])
    except (ValueError, IndexError):
        return 0.0


def params_from_dict(p: Dict[str, str], imax: int, jmax: int, dx: float,
                     nthreads: int = 1, grain_d=None) -> orc_params:
    """Parameter semantics of LSDCatchmentModel::initialise_variables
    (LSDCatchmentModel.cpp:842-1610) for the keys the synthetic workloads use,
    plus the max_time_step cap of initialise_arrays (1680-1683)."""
    g = lambda k, d: p.get(k, d)
    o = orc_params()
    o.imax, o.jmax, o.DX = imax, jmax, dx
    o.courant_number = _atof(g("courant_number", "0.7"))
    o.froude_limit = _atof(g("froude_num_limit", "0.8"))
    # water_depth_erosion_threshold sets hflow_threshold (LCM.cpp:1318-1323);
    # a later key wins, as in the reference's if/else chain over file lines.
    hthr = 0.00001
    for k, v in p.items():
        if k in ("hflow_threshold", "water_depth_erosion_threshold"):
            hthr = _atof(v)
    o.hflow_threshold = hthr
    o.edgeslope = _atof(g("slope_on_edge_cell", "0.001"))
    o.mannings = _atof(g("mannings_n", "0.04"))
    o.spatially_var_mannings = 1 if g("spatially_variable_mannings_on", "no") == "yes" else 0
    o.water_depth_erosion_threshold = 0.01
    o.ERODEFACTOR = _atof(g("erosion_limit", "0.05"))
    o.active = _atof(g("active_layer_thickness", "0.2"))
    o.Csuspmax = 0.05
    o.max_vel = _atof(g("max_tau_velocity", "5"))
    law = g("transport_law", "")
    o.wilcock = 1 if law == "wilcock" else 0
    o.einstein = 1 if law == "einstein" else 0
    d = list(REF_D) if grain_d is None else [0.0] + list(grain_d)
    for i in range(10):
        o.d[i] = d[i]
    for i in range(11):
        o.dprop[i] = DPROP[i]
        o.fallVelocity[i] = FALL_VELOCITY[i]
    spatial_rain = g("spatial_var_rain", "no") == "yes"
    o.rfnum = _atoi(g("num_unique_rain_cells", "1")) if spatial_rain else 1
    o.M = _atof(g("topmodel_m_value", "0.005"))
    rdt = _atof(g("rain_data_time_step", "60"))
    if rdt < 1:
        rdt = 1
    o.rain_data_time_step = rdt
    mts = _atoi(g("max_time_step", "0"))
    if mts // 60 > rdt:
        mts = int(rdt) * 60
    o.max_time_step = mts
    o.in_out_difference_allowed = _atof(g("in_out_difference", "0"))
    o.MIN_Q = _atof(g("min_q_for_depth_calc", "0.01"))
    o.MIN_Q_MAXVAL = _atof(g("max_q_for_depth_calc", "1000.0"))
    o.maxcycle = _atoi(g("max_run_duration", "1000"))
    o.output_file_save_interval = float(_atoi(g("timeseries_save_interval", "60")))
    o.hydro_only = 1 if g("hydro_model_only", "no") == "yes" else 0
    o.bedrock_erosion_threshold = _atof(g("bedrock_erosion_threshold", "0"))
    o.bedrock_erosion_rate = 0.0
    o.lateral_constant = _atof(g("lateral_erosion_const", "0.0000002"))
    o.chann_lateral_erosion = _atof(g("chann_lateral_erosion", "20"))
    o.veg_lat_restriction = _atof(g("veg_erosion_prop", "0.1"))
    o.vegTauCrit = _atof(g("vegetation_crit_shear", "100"))
    o.scan_interval = 5
    o.nthreads = nthreads
    o.spatially_complex_rainfall = 1 if g("spatially_complex_rainfall_on", "no") == "yes" else 0
    o.no_data_value = -9999.0  # the synthetic DEMs' NODATA_value header
    o.reach_mode = 1 if g("reach_mode", "no") == "yes" else 0
    o.div_inputs = int(_atof(g("divide_inputs_by", "1")))
    o.reach_input_data_timestep = _atoi(g("reach_input_data_timestep", "60"))
    o.creep_rate = _atof(g("creep_rate", "0.0025"))
    o.failureangle = _atof(g("slope_failure_thresh", "45"))
    return o


def parse_params(path: str) -> Dict[str, str]:
    """``key: value # comment`` lines, keys lower-cased, later keys win
    (LSDUtils::parse_line, LSDUtils.cpp:12-77; LSDCatchmentModel.cpp:858-867)."""
    out: Dict[str, str] = {}
    with open(path) as f:
        for line in f:
            line = line.split("#", 1)[0]
            if ":" not in line:
                continue
            k, v = line.split(":", 1)
            k = k.strip().lower()
            v = v.strip().split()
            if k:
                out[k] = v[0] if v else ""
    return out


class FileWorkload:
    """A reference-style input directory (params + DEM + rain) seen as a
    Workload for OracleModel.from_workload."""

    def __init__(self, path: str, params_file: str):
        from hail_caesar_amd.inputs import read_asc
        self.path = path
        self.params_file = params_file
        self.params = parse_params(os.path.join(path, params_file))
        self.params.setdefault("read_path", "./")
        dem_path = os.path.join(path, self.params["read_fname"] + "." + self.params.get("dem_read_extension", "asc"))
        with open(dem_path) as f:
            hdr = [f.readline().split() for _ in range(6)]
        self.dx = float(hdr[4][1])
        self.dem = read_asc(dem_path)
        self.mannings = None
        self.water_init = None
        self.extra = {}
        if self.params.get("spatially_variable_mannings_on") == "yes":
            self.mannings = read_asc(os.path.join(path, self.params["spatial_mannings_dem_file"]))
        if self.params.get("water_init_from_raster_on") == "yes":
            self.water_init = read_asc(os.path.join(path, self.params["water_init_raster_file"]))


def pad(a: np.ndarray, fill: float) -> np.ndarray:
    out = np.full((a.shape[0] + 2, a.shape[1] + 2), fill, dtype=a.dtype)
    out[1:-1, 1:-1] = a
    return out


class OracleModel:
    """FP64 CPU restatement of the hot path, driven from a Workload."""

    _lib = None

    @classmethod
    def lib(cls):
        if cls._lib is None:
            if not os.path.exists(ORACLE_SO):
                raise RuntimeError(f"oracle library missing: {ORACLE_SO} (run __graft_entry__.build())")
            L = ctypes.CDLL(ORACLE_SO)
            L.orc_create.restype = ctypes.c_void_p
            L.orc_create.argtypes = [ctypes.POINTER(orc_params)]
            L.orc_destroy.argtypes = [ctypes.c_void_p]
            L.orc_field.restype = ctypes.POINTER(ctypes.c_double)
            L.orc_field.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_long)]
            L.orc_field_int.restype = ctypes.POINTER(ctypes.c_int)
            L.orc_field_int.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_long)]
            L.orc_set_rain.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
            L.orc_set_reach.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
            L.orc_reach_input.argtypes = [ctypes.c_void_p]
            L.orc_scalar.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int]
            L.orc_initialise.argtypes = [ctypes.c_void_p]
            for fn in ("orc_flow_route", "orc_scan_area", "orc_set_maximum_timestep",
                       "orc_call_erosion", "orc_drainage_area"):
                getattr(L, fn).argtypes = [ctypes.c_void_p]
            for fn in ("orc_depth_update", "orc_water_flux_out", "orc_get_flow_timestep"):
                getattr(L, fn).argtypes = [ctypes.c_void_p]
                getattr(L, fn).restype = ctypes.c_double
            L.orc_global_landsliding.argtypes = [ctypes.c_void_p, ctypes.c_double]
            L.orc_global_landsliding.restype = ctypes.c_int
            L.orc_creep.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
            L.orc_in_channel_landsliding.argtypes = [ctypes.c_void_p, ctypes.c_double]
            L.orc_erode.argtypes = [ctypes.c_void_p, ctypes.c_double]
            L.orc_erode.restype = ctypes.c_double
            L.orc_input_points.argtypes = [ctypes.c_void_p]
            L.orc_grain_call.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_double, ctypes.c_int, ctypes.c_int]
            L.orc_grain_call.restype = ctypes.c_double
            L.orc_iterate.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_void_p]
            L.orc_series.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
            cls._lib = L
        return cls._lib

    def __init__(self, params: orc_params):
        self.L = self.lib()
        self.p = params
        self.h = self.L.orc_create(ctypes.byref(params))
        self.shape = (params.imax + 2, params.jmax + 2)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def field(self, name: str) -> np.ndarray:
        n = ctypes.c_long()
        ptr = self.L.orc_field(self.h, name.encode(), ctypes.byref(n))
        if not ptr:
            raise KeyError(name)
        arr = np.ctypeslib.as_array(ptr, shape=(n.value,))
        ncell = self.shape[0] * self.shape[1]
        k = n.value // ncell
        return arr.reshape(self.shape) if k == 1 else arr.reshape(self.shape + (k,))

    def field_int(self, name: str) -> np.ndarray:
        n = ctypes.c_long()
        ptr = self.L.orc_field_int(self.h, name.encode(), ctypes.byref(n))
        if not ptr:
            raise KeyError(name)
        return np.ctypeslib.as_array(ptr, shape=(n.value,)).reshape(self.shape)

    def scalar(self, name: str, value: Optional[float] = None) -> float:
        v = ctypes.c_double(0.0 if value is None else value)
        rc = self.L.orc_scalar(self.h, name.encode(), ctypes.byref(v), 0 if value is None else 1)
        if rc != 0:
            raise KeyError(name)
        return v.value

    def set_rain(self, rain: np.ndarray) -> None:
        r = np.ascontiguousarray(rain, dtype=np.float64)
        if r.ndim == 1:
            r = r[:, None]
        self.L.orc_set_rain(self.h, r.ctypes.data, r.shape[0], r.shape[1])

    @classmethod
    def from_workload(cls, w, nthreads: int = 1, grain_d=None, initialise: bool = True):
        from hail_caesar_amd import inputs  # noqa: F401  (file formats)
        dem = w.dem
        imax, jmax = dem.shape
        p = params_from_dict(w.params, imax, jmax, w.dx, nthreads,
                             grain_d=grain_d if grain_d is not None else w.extra.get("grain_d"))
        m = cls(p)
        m.field("elev")[...] = pad(dem, -9999.0)
        if w.mannings is not None and p.spatially_var_mannings:
            m.field("mannings")[1:-1, 1:-1] = w.mannings
        if w.water_init is not None:
            m.field("water_depth")[1:-1, 1:-1] = w.water_init
        hydro = os.path.join(w.path, "hydroindex.asc")
        if w.params.get("spatial_var_rain", "no") == "yes":
            from hail_caesar_amd.inputs import read_asc
            m.field_int("rfarea")[1:-1, 1:-1] = read_asc(hydro).astype(np.int32)
        rain = np.loadtxt(os.path.join(w.path, w.params["rainfall_data_file"]), ndmin=2)
        m.set_rain(rain)
        if not p.hydro_only and w.params.get("read_in_graindata_from_file", "no") == "yes":
            m.load_graindata(os.path.join(w.path, w.params["grain_data_file"]))
        if p.reach_mode:
            m.load_reach(w)
        if initialise:
            m.L.orc_initialise(m.h)
        return m

    def load_reach(self, w) -> None:
        """Reach input points and files (LSDCatchmentModel.cpp:1766-1845,
        read_reachfile 718-760)."""
        for k in range(3):
            if w.params.get(f"reach{k + 1}_input_on", "no") != "yes":
                continue
            rows = np.loadtxt(os.path.join(w.path, w.params[f"reach{k + 1}_input_file"]), ndmin=2)
            rows = np.ascontiguousarray(rows, dtype=np.float64)
            self.L.orc_set_reach(self.h, k, _atoi(w.params[f"reach{k + 1}_x"]),
                                 _atoi(w.params[f"reach{k + 1}_y"]), rows.ctypes.data,
                                 rows.shape[0], rows.shape[1])

    def load_graindata(self, path: str) -> None:
        """Grain data file semantics of ingest_graindata_from_file
        (LSDCatchmentModel.cpp:611-683)."""
        g = self.field("grain")
        st = self.field("strata").reshape(self.shape + (10, 9))
        hg = self.field_int("hasgrain")
        gi = self.field_int("index")
        imax, jmax = self.p.imax, self.p.jmax
        data = np.loadtxt(path, ndmin=2)
        tot = 0
        for row in data:
            tot += 1
            x1 = min(int(row[0]), imax)
            y1 = min(int(row[1]), jmax)
            hg[x1, y1] = 1
            gi[x1, y1] = tot  # index[x1][y1] = grain_array_tot (LCM.cpp:641-656)
            g[x1, y1, :] = row[3:14]
            st[x1, y1] = row[14:104].reshape(10, 9)
        self.scalar("grain_array_tot", float(tot))

    def load_grain_arrays(self, graindata: np.ndarray) -> None:
        """The grain data file's values given per cell in memory ((imax, jmax,
        101): grain[0..10] then strata[z][n]), every cell with a row numbered
        in row-major order, as a file of one line per cell would number them
        (LSDCatchmentModel.cpp:611-683)."""
        imax, jmax = self.p.imax, self.p.jmax
        gd = np.asarray(graindata, dtype=np.float64).reshape(imax, jmax, 101)
        self.field("grain")[1:-1, 1:-1, :] = gd[:, :, :11]
        self.field("strata").reshape(self.shape + (90,))[1:-1, 1:-1, :] = gd[:, :, 11:]
        self.field_int("hasgrain")[1:-1, 1:-1] = 1
        self.field_int("index")[1:-1, 1:-1] = np.arange(1, imax * jmax + 1, dtype=np.int32).reshape(imax, jmax)
        self.scalar("grain_array_tot", float(imax * jmax))

    def load_snapshot(self, snap: dict) -> None:
        """Inverse of RefModel.snapshot for the oracle's dense layout."""
        names = {"spat_var_mannings": "mannings"}
        for k in RefModel.SNAP_2D:
            if k == "spat_var_mannings" and not self.p.spatially_var_mannings:
                continue
            self.field(names.get(k, k))[...] = snap[k]
        self.field_int("mask")[...] = snap["mask"]
        self.field("vel_dir")[...] = snap["vel_dir"]
        for k in ("sr", "sl", "su", "sd"):
            self.field(k)[...] = snap[k]
        self.field("grain")[...] = snap["grain"]
        self.field("strata")[...] = snap["strata"].reshape(self.shape + (90,))
        self.field_int("hasgrain")[...] = snap["hasgrain"]
        if "index" in snap:
            self.field_int("index")[...] = snap["index"]
            self.scalar("grain_array_tot", float(snap["grain_array_tot"]))
        for k in ("time_step", "maxdepth", "erode_mult", "globalsediq", "counter", "cycle"):
            self.scalar(k, float(snap[k]))

    def iterate(self, n: int):
        tt = np.zeros(n)
        cy = np.zeros(n)
        dt = np.zeros(n)
        k = self.L.orc_iterate(self.h, n, tt.ctypes.data, cy.ctypes.data, dt.ctypes.data)
        return tt[:k], cy[:k], dt[:k]

    def series(self) -> np.ndarray:
        rows = self.L.orc_series(self.h, None, 0)
        out = np.zeros((rows, 14))
        if rows:
            self.L.orc_series(self.h, out.ctypes.data, rows)
        return out

    def call(self, name: str, *args):
        fn = getattr(self.L, "orc_" + name)
        return fn(self.h, *args)


class RefModel:
    """The reference model itself (compiled from its sources by oracle/Makefile)."""

    _lib = None

    @staticmethod
    def available() -> bool:
        return os.path.exists(REF_SO)

    @classmethod
    def lib(cls):
        if cls._lib is None:
            L = ctypes.CDLL(REF_SO)
            L.ref_create.restype = ctypes.c_void_p
            L.ref_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
            L.ref_destroy.argtypes = [ctypes.c_void_p]
            L.ref_iterate.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_void_p]
            L.ref_call.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
            L.ref_call.restype = ctypes.c_double
            L.ref_grain_call.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_double, ctypes.c_int, ctypes.c_int]
            L.ref_grain_call.restype = ctypes.c_double
            for fn in ("ref_field",):
                getattr(L, fn).restype = ctypes.POINTER(ctypes.c_double)
                getattr(L, fn).argtypes = [ctypes.c_void_p, ctypes.c_char_p,
                                           ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
            L.ref_field_int.restype = ctypes.POINTER(ctypes.c_int)
            L.ref_field_int.argtypes = [ctypes.c_void_p, ctypes.c_char_p,
                                        ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
            L.ref_field3.restype = ctypes.POINTER(ctypes.c_double)
            L.ref_field3.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int),
                                     ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
            L.ref_vec.restype = ctypes.POINTER(ctypes.c_double)
            L.ref_vec.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]
            L.ref_vec_int.restype = ctypes.POINTER(ctypes.c_int)
            L.ref_vec_int.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]
            L.ref_scalar.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int]
            L.ref_set_threads.argtypes = [ctypes.c_int]
            cls._lib = L
        return cls._lib

    CALLS = {"flow_route": 0, "depth_update": 1, "scan_area": 2, "water_flux_out": 3,
             "set_maximum_timestep": 4, "increment_counters": 5, "catchment_waterinputs": 6,
             "call_erosion": 7, "set_loop_cycle": 8, "set_inputoutput_diff": 9,
             "write_output_timeseries": 10, "drainage_area": 11, "input_points": 12,
             "erode": 13, "get_flow_timestep": 14, "topmodel_runoff": 15,
             "global_landsliding": 16, "creep": 17,
             "in_channel_landsliding": 18, "reach_water_and_sediment_input": 19}

    def __init__(self, path: str, params_file: str, threads: int = 1):
        self.L = self.lib()
        self.L.ref_set_threads(threads)
        cwd = os.getcwd()
        os.chdir(path)
        try:
            self.h = self.L.ref_create(b"./", params_file.encode())
        finally:
            os.chdir(cwd)
        self.path = path

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ref_destroy(self.h)
            self.h = None

    def field(self, name: str) -> np.ndarray:
        d0, d1 = ctypes.c_int(), ctypes.c_int()
        ptr = self.L.ref_field(self.h, name.encode(), ctypes.byref(d0), ctypes.byref(d1))
        if not ptr:
            raise KeyError(name)
        return np.ctypeslib.as_array(ptr, shape=(d0.value, d1.value))

    def field_int(self, name: str) -> np.ndarray:
        d0, d1 = ctypes.c_int(), ctypes.c_int()
        ptr = self.L.ref_field_int(self.h, name.encode(), ctypes.byref(d0), ctypes.byref(d1))
        if not ptr:
            raise KeyError(name)
        return np.ctypeslib.as_array(ptr, shape=(d0.value, d1.value))

    def field3(self, name: str) -> np.ndarray:
        d0, d1, d2 = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        ptr = self.L.ref_field3(self.h, name.encode(), ctypes.byref(d0), ctypes.byref(d1), ctypes.byref(d2))
        if not ptr:
            raise KeyError(name)
        return np.ctypeslib.as_array(ptr, shape=(d0.value, d1.value, d2.value))

    def vec(self, name: str) -> np.ndarray:
        n = ctypes.c_int()
        ptr = self.L.ref_vec(self.h, name.encode(), ctypes.byref(n))
        return np.ctypeslib.as_array(ptr, shape=(n.value,))

    def scalar(self, name: str, value: Optional[float] = None) -> float:
        v = ctypes.c_double(0.0 if value is None else value)
        rc = self.L.ref_scalar(self.h, name.encode(), ctypes.byref(v), 0 if value is None else 1)
        if rc != 0:
            raise KeyError(name)
        return v.value

    def mask(self) -> np.ndarray:
        """down_scan lists converted to a padded 0/1 mask."""
        ds = self.field_int("down_scan")  # (jmax+2, imax+2)
        imax = int(self.scalar("imax"))
        jmax = int(self.scalar("jmax"))
        m = np.zeros((imax + 2, jmax + 2), dtype=np.int32)
        for y in range(1, jmax + 1):
            row = ds[y, 1:]
            k = np.argmax(row == 0) if np.any(row == 0) else len(row)
            m[row[:k], y] = 1
        return m

    SNAP_2D = ("elev", "water_depth", "qx", "qy", "qxs", "qys", "Vsusptot", "bedrock", "Tau",
               "Vel", "ss", "spat_var_mannings")

    def snapshot(self) -> dict:
        """Full erosion-relevant state with the lazily indexed grain rows
        (index[x][y] -> grain/strata row, LSDCatchmentModel.cpp:3949-3983)
        expanded to one dense block per cell."""
        out = {k: self.field(k).copy() for k in self.SNAP_2D}
        out["mask"] = self.mask()
        out["vel_dir"] = self.field3("vel_dir").copy()
        for k in ("sr", "sl", "su", "sd"):
            out[k] = self.field3(k).copy()
        idx = self.field_int("index")
        g = self.field("grain")
        st = self.field3("strata")
        has = idx != -9999
        ni, nj = idx.shape
        grain = np.zeros((ni, nj, 11))
        strata = np.zeros((ni, nj, 10, 9))
        grain[has] = g[idx[has]]
        strata[has] = st[idx[has]][:, :, :9]
        out["grain"] = grain
        out["strata"] = strata
        out["hasgrain"] = has.astype(np.int32)
        out["index"] = idx.copy()
        out["grain_array_tot"] = self.scalar("grain_array_tot")
        for k in ("time_step", "maxdepth", "erode_mult", "globalsediq", "counter", "cycle"):
            out[k] = self.scalar(k)
        out["sum_grain"] = self.vec("sum_grain").copy()
        return out

    def call(self, name: str, arg: float = 0.0) -> float:
        cwd = os.getcwd()
        os.chdir(self.path)
        try:
            return self.L.ref_call(self.h, self.CALLS[name], arg)
        finally:
            os.chdir(cwd)

    def iterate(self, n: int):
        tt = np.zeros(n)
        cy = np.zeros(n)
        dt = np.zeros(n)
        cwd = os.getcwd()
        os.chdir(self.path)
        try:
            k = self.L.ref_iterate(self.h, n, tt.ctypes.data, cy.ctypes.data, dt.ctypes.data)
        finally:
            os.chdir(cwd)
        return tt[:k], cy[:k], dt[:k]

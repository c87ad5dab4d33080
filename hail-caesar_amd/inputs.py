"""Seeded synthetic inputs in the reference's own file formats.

Writes ASCII DEMs (``ncols/nrows/xllcorner/yllcorner/cellsize/NODATA_value``
header as read by ``LSDRaster::read_ascii_raster``, src/topotools/LSDRaster.cpp:797-852),
rainfall files (one row per ``rain_data_time_step``, one column per rain cell,
read by ``LSDCatchmentModel::read_rainfalldata``, LSDCatchmentModel.cpp:694-717),
grain data files (``x y idx grain[0..10] strata[10][9]`` per line,
LSDCatchmentModel.cpp:611-683) and ``key: value`` parameter files
(src/catchmentmodel/LSDUtils.cpp:12-77).

Elevations are written with ``%.17g`` so the reference and this build read
bit-identical FP64 values.  NODATA is the literal -9999 that every nodata test
in the reference compares against (LSDCatchmentModel.cpp:2853, 3086).
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

NODATA = -9999.0

# Grain sizes hard-coded in the reference (LSDCatchmentModel.hpp:557-565).
REF_GRAIN_D = [0.000065, 0.001, 0.002, 0.004, 0.008, 0.016, 0.032, 0.064, 0.128]
# Nine-fraction set of the flow+erosion workload (d1 overridden to 0.0005 m).
EROSION_GRAIN_D = [0.0005, 0.001, 0.002, 0.004, 0.008, 0.016, 0.032, 0.064, 0.128]


# ---------------------------------------------------------------------------
# terrain
# ---------------------------------------------------------------------------
def diamond_square(p: int, roughness: float, rng: np.random.Generator) -> np.ndarray:
    """Diamond-square fractal on a (2**p+1)^2 grid; the displacement scale is
    multiplied by ``roughness`` at every level."""
    n = 2 ** p + 1
    a = np.zeros((n, n), dtype=np.float64)
    a[0, 0], a[0, -1], a[-1, 0], a[-1, -1] = rng.uniform(-1.0, 1.0, 4)
    step = n - 1
    scale = 1.0
    while step > 1:
        half = step // 2
        # diamond step: centres of squares
        c = (a[0:-1:step, 0:-1:step] + a[0:-1:step, step::step]
             + a[step::step, 0:-1:step] + a[step::step, step::step]) * 0.25
        a[half:-1:step, half:-1:step] = c + rng.uniform(-scale, scale, c.shape)
        # square step: edge midpoints.  Rows with even offset (midpoints of
        # horizontal edges) and odd offset (vertical edges).
        for (r0, c0) in ((0, half), (half, 0)):
            rows = np.arange(r0, n, step)
            cols = np.arange(c0, n, step)
            rr, cc = np.meshgrid(rows, cols, indexing="ij")
            tot = np.zeros(rr.shape)
            cnt = np.zeros(rr.shape)
            for dr, dc in ((-half, 0), (half, 0), (0, -half), (0, half)):
                r2, c2 = rr + dr, cc + dc
                ok = (r2 >= 0) & (r2 < n) & (c2 >= 0) & (c2 < n)
                tot[ok] += a[r2[ok], c2[ok]]
                cnt[ok] += 1
            a[rr, cc] = tot / cnt + rng.uniform(-scale, scale, rr.shape)
        step = half
        scale *= roughness
    return a


def fractal_noise(nrows: int, ncols: int, amplitude: float, roughness: float,
                  rng: np.random.Generator) -> np.ndarray:
    """Zero-mean fractal field with peak-to-peak range ``amplitude``."""
    p = max(1, int(math.ceil(math.log2(max(nrows, ncols) - 1))) if max(nrows, ncols) > 2 else 1)
    ds = diamond_square(p, roughness, rng)[:nrows, :ncols]
    ds = ds - ds.mean()
    rng_pp = ds.max() - ds.min()
    if rng_pp > 0:
        ds *= amplitude / rng_pp
    return ds


def carve_channel(z: np.ndarray, depth: float, width: int = 3) -> np.ndarray:
    """Carve a ``width``-cell channel ``depth`` metres deep down the central
    column to the south edge (row index increases southward)."""
    z = z.copy()
    nrows, ncols = z.shape
    c0 = ncols // 2 - width // 2
    z[:, c0:c0 + width] -= depth
    return z


def carve_d8_network(z: np.ndarray, block: int = 64, min_acc: int = 6, depth: float = 3.0,
                     bed_drop: float = 0.002) -> np.ndarray:
    """Carve a D8-consistent drainage network draining to the south edge.

    A coarse grid of ``block`` x ``block`` cell means is priority-flooded from
    its south row, giving every coarse cell a D8 receiver; coarse cells with a
    flow accumulation of at least ``min_acc`` form the network.  Each network
    link is rasterised as a D8 path between block centres (the outlets run
    straight to the south edge) and given a bed that falls strictly
    downstream: at least ``bed_drop`` metres per cell and at least ``depth``
    below the lowest terrain the link crosses.  Path cells take the bed
    elevation and their 8-neighbours are lowered to at most bed + depth/2, so
    every path cell has a strictly lower D8 neighbour on the network, down to
    the south edge.  NODATA cells are left alone.  Returns (z, channel mask)."""
    import heapq
    z = z.copy()
    nr, nc = z.shape
    gr, gc = max(1, nr // block), max(1, nc // block)
    valid = z > NODATA
    zf = np.where(valid, z, np.nan)
    zc = np.nanmean(zf[:gr * block, :gc * block].reshape(gr, block, gc, block), axis=(1, 3))
    zc = np.where(np.isnan(zc), np.nanmax(zc), zc)
    recv = -np.ones(gr * gc, dtype=np.int64)
    seen = np.zeros(gr * gc, dtype=bool)
    heap = []
    for j in range(gc):  # the south row drains off the grid
        i = (gr - 1) * gc + j
        seen[i] = True
        heapq.heappush(heap, (float(zc[gr - 1, j]), i))
    order = []
    while heap:
        e, i = heapq.heappop(heap)
        order.append(i)
        r, c = divmod(i, gc)
        for dr in (-1, 0, 1):
            for dc in (-1, 0, 1):
                rr, cc = r + dr, c + dc
                if (dr or dc) and 0 <= rr < gr and 0 <= cc < gc:
                    k = rr * gc + cc
                    if not seen[k]:
                        seen[k] = True
                        recv[k] = i
                        heapq.heappush(heap, (max(float(zc[rr, cc]), e), k))
    acc = np.ones(gr * gc)
    for i in reversed(order):
        if recv[i] >= 0:
            acc[recv[i]] += acc[i]
    centre = lambda i: (divmod(i, gc)[0] * block + block // 2, divmod(i, gc)[1] * block + block // 2)

    def path(a, b):
        (r, c), (r1, c1) = a, b
        out = [(r, c)]
        while (r, c) != (r1, c1):
            r += int(np.sign(r1 - r))
            c += int(np.sign(c1 - c))
            out.append((r, c))
        return out

    bed = {}
    links = []
    for i in order:  # outlet first: a link's receiver has its bed already
        if acc[i] < min_acc:
            continue
        a = centre(i)
        if recv[i] < 0:
            pts = path(a, (nr - 1, a[1]))
            end = None
        else:
            pts = path(a, centre(recv[i]))
            end = bed.get(int(recv[i]))
        terr = [z[p] for p in pts if valid[p]]
        low = (min(terr) if terr else 0.0) - depth
        n = len(pts) - 1
        b_end = end if end is not None else low
        b_top = max(low, b_end + bed_drop * max(n, 1))
        bed[int(i)] = b_top
        if end is not None:
            pts = pts[:-1]  # the receiver's own cell keeps its bed
        links.append([(r, c, b_top + (b_end - b_top) * (k / max(n, 1))) for k, (r, c) in enumerate(pts)
                      if valid[r, c]])
    # banks first, then the beds: a cell on several links takes the lowest
    # bed, so from every path cell some link continues strictly downhill
    for pts in links:
        for r, c, zb in pts:
            for dr in (-1, 0, 1):
                for dc in (-1, 0, 1):
                    rr, cc = r + dr, c + dc
                    if (dr or dc) and 0 <= rr < nr and 0 <= cc < nc and valid[rr, cc]:
                        z[rr, cc] = min(z[rr, cc], max(zb + depth / 2, z[rr, cc] - depth))
    for pts in links:
        for r, c, zb in pts:
            z[r, c] = zb
    chan = np.zeros(z.shape, dtype=bool)
    for pts in links:
        for r, c, zb in pts:
            chan[r, c] = True
    for pts in links:
        for r, c, zb in pts:
            z[r, c] = min(z[r, c], zb)
    return z, chan


def planar_dem(nrows: int, ncols: int, dx: float, slope: float, amplitude: float,
               roughness: float, seed: int, channel_depth: float = 2.0,
               south_only: bool = True, base: float = 10.0) -> np.ndarray:
    """Planar slope toward the south edge plus fractal noise and a carved
    channel.  With ``south_only`` the north/east/west edge ring is NODATA so
    that only the south edge can drain (the reference drains all four edges,
    LSDCatchmentModel.cpp:3005-3047)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    i = np.arange(nrows, dtype=np.float64)[:, None]
    z = base + slope * dx * (nrows - i) + fractal_noise(nrows, ncols, amplitude, roughness, rng)
    z = carve_channel(z, channel_depth)
    if south_only:
        z[0, :] = NODATA
        z[:, 0] = NODATA
        z[:, -1] = NODATA
    return z


# ---------------------------------------------------------------------------
# file writers
# ---------------------------------------------------------------------------
def write_asc(path: str, arr: np.ndarray, dx: float, nodata: float = NODATA,
              fmt: str = "%.17g") -> None:
    nrows, ncols = arr.shape
    with open(path, "w") as f:
        f.write(f"ncols {ncols}\nnrows {nrows}\nxllcorner 0\nyllcorner 0\n")
        f.write(f"cellsize {dx:.17g}\nNODATA_value {nodata:.17g}\n")
        np.savetxt(f, arr, fmt=fmt, delimiter=" ")


def write_asc_int(path: str, arr: np.ndarray, dx: float, nodata: int = -9999) -> None:
    nrows, ncols = arr.shape
    with open(path, "w") as f:
        f.write(f"ncols {ncols}\nnrows {nrows}\nxllcorner 0\nyllcorner 0\n")
        f.write(f"cellsize {dx:.17g}\nNODATA_value {nodata}\n")
        np.savetxt(f, arr.astype(np.int64), fmt="%d", delimiter=" ")


def read_asc(path: str) -> np.ndarray:
    with open(path) as f:
        hdr = [f.readline() for _ in range(6)]
    ncols = int(hdr[0].split()[1])
    nrows = int(hdr[1].split()[1])
    a = np.loadtxt(path, skiprows=6, ndmin=2)
    return a.reshape(nrows, ncols)


def write_rain(path: str, rates_mm_h: np.ndarray, pad_rows: int = 200) -> None:
    """One row per rain data step, one column per rain cell (mm/h).  Zero rows
    are appended because the reference indexes past the end of the storm
    (LSDCatchmentModel.cpp:3676)."""
    # values are made float32-exact so any reader (the reference parses them
    # as float) recovers identical numbers from the 9-digit text
    r = np.asarray(rates_mm_h, dtype=np.float32).astype(np.float64)
    if r.ndim == 1:
        r = r[:, None]
    r = np.vstack([r, np.zeros((pad_rows, r.shape[1]))])
    np.savetxt(path, r, fmt="%.9g", delimiter=" ")


def write_graindata(path: str, props: np.ndarray, active: float,
                    strata_props: Optional[np.ndarray] = None,
                    cells: Optional[np.ndarray] = None) -> None:
    """Grain data file: one line per cell, ``x y idx grain[0..10]
    strata[z=0..9][n=0..8]`` with 1-based padded coordinates x=row, y=col
    (LSDCatchmentModel.cpp:611-683).  ``props`` is (nrows, ncols, 9) surface
    proportions; grain[n] = active*props[n-1] for n=1..9, grain[0]=grain[10]=0.
    Strata layers get the same proportions times ``active``.  ``cells`` (k x 2,
    0-based) lists the cells that get a line, in line order; default every
    cell in row-major order."""
    nrows, ncols, nf = props.shape
    assert nf == 9
    if strata_props is None:
        strata_props = props
    if cells is None:
        cells = np.argwhere(np.ones((nrows, ncols), dtype=bool))
    lines: List[str] = []
    idx = 0
    for i, j in np.asarray(cells):
        idx += 1
        g = [0.0] + list(active * props[i, j]) + [0.0]
        st = list(np.tile(active * strata_props[i, j], 10))
        vals = [str(i + 1), str(j + 1), str(idx)] + ["%.17g" % v for v in g + st]
        lines.append(" ".join(vals))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def graindata_array(props: np.ndarray, active: float) -> np.ndarray:
    """The values of write_graindata's lines without their x y index columns:
    (nrows, ncols, 101) = grain[0..10] (active*props for fractions 1..9) and
    strata[z=0..9][n=0..8] (active*props in every layer), for in-memory runs."""
    nrows, ncols, nf = props.shape
    assert nf == 9
    out = np.zeros((nrows, ncols, 101))
    out[:, :, 1:10] = active * props
    out[:, :, 11:] = np.tile(active * props, (1, 1, 10))
    return out


# Parameter defaults shared by every synthetic workload.  Keys and meanings
# follow LSDCatchmentModel::initialise_variables (LSDCatchmentModel.cpp:842-1610).
BASE_PARAMS: Dict[str, str] = {
    "dem_read_extension": "asc",
    "dem_write_extension": "asc",
    "timeseries_save_interval": "5",
    "min_time_step": "0",
    "max_time_step": "300",
    "run_time_start": "0",
    "memory_limit": "1",
    "transport_law": "wilcock",
    "max_tau_velocity": "5",
    "active_layer_thickness": "0.1",
    "chann_lateral_erosion": "20",
    "erosion_limit": "0.02",
    "suspended_sediment_on": "yes",
    "read_in_graindata_from_file": "no",
    "bedrock_layer_on": "no",
    "lateral_erosion_on": "no",
    "hydro_model_only": "yes",
    "topmodel_m_value": "0.0005",
    "in_out_difference": "0",
    "min_q_for_depth_calc": "0.00001",
    "max_q_for_depth_calc": "1000.0",
    "slope_on_edge_cell": "0.01",
    "evaporation_rate": "0.0",
    "courant_number": "0.7",
    "froude_num_limit": "0.8",
    "mannings_n": "0.04",
    "hflow_threshold": "0.001",
    "rainfall_data_on": "yes",
    "rain_data_time_step": "60",
    "spatial_var_rain": "no",
    "num_unique_rain_cells": "1",
    "spatially_complex_rainfall_on": "no",
    "vegetation_on": "no",
    "grass_grow_rate": "0.0",
    "vegetation_crit_shear": "0.0",
    "veg_erosion_prop": "0.0",
    "raster_output_interval": "100000",
    "write_waterdepth_file": "no",
    "write_elev_file": "no",
    "write_grainsize_file": "no",
    "write_elevdiff_file": "no",
    "debug_print_cycle": "no",
}


def write_params(path: str, params: Dict[str, str]) -> None:
    with open(path, "w") as f:
        f.write("# synthetic workload parameter file\n")
        for k, v in params.items():
            f.write(f"{k}: {v}\n")


@dataclass
class Workload:
    """A complete on-disk input set: directory, parameter file name, and the
    arrays that were written (unpadded, row 0 = north)."""
    path: str
    params_file: str
    params: Dict[str, str]
    dem: np.ndarray
    dx: float
    mannings: Optional[np.ndarray] = None
    water_init: Optional[np.ndarray] = None
    rain: Optional[np.ndarray] = None
    extra: Dict[str, object] = field(default_factory=dict)


def make_workload(path: str, dem: np.ndarray, dx: float, rain_mm_h: np.ndarray,
                  rain_step_min: float, run_hours: int, overrides: Dict[str, str] = None,
                  mannings: Optional[np.ndarray] = None,
                  water_init: Optional[np.ndarray] = None,
                  hydroindex: Optional[np.ndarray] = None,
                  graindata: Optional[np.ndarray] = None,
                  grain_cells: Optional[np.ndarray] = None,
                  name: str = "run") -> Workload:
    os.makedirs(path, exist_ok=True)
    p = dict(BASE_PARAMS)
    p["read_path"] = "./"
    p["write_path"] = "./out/"
    p["read_fname"] = "dem"
    p["write_fname"] = f"{name}.dat"
    p["rainfall_data_file"] = "rain.txt"
    p["rain_data_time_step"] = "%g" % rain_step_min
    p["max_run_duration"] = str(int(run_hours))
    os.makedirs(os.path.join(path, "out"), exist_ok=True)
    write_asc(os.path.join(path, "dem.asc"), dem, dx)
    write_rain(os.path.join(path, "rain.txt"), rain_mm_h)
    if mannings is not None:
        write_asc(os.path.join(path, "mannings.asc"), mannings, dx)
        p["spatially_variable_mannings_on"] = "yes"
        p["spatial_mannings_dem_file"] = "mannings.asc"
    if water_init is not None:
        write_asc(os.path.join(path, "water_init.asc"), water_init, dx)
        p["water_init_from_raster_on"] = "yes"
        p["water_init_raster_file"] = "water_init.asc"
    if hydroindex is not None:
        write_asc_int(os.path.join(path, "hydroindex.asc"), hydroindex, dx)
        p["spatial_var_rain"] = "yes"
        p["hydroindex_file"] = "hydroindex.asc"
        p["num_unique_rain_cells"] = str(int(hydroindex.max()))
    if graindata is not None:
        active = float((overrides or {}).get("active_layer_thickness", p["active_layer_thickness"]))
        write_graindata(os.path.join(path, "grain.dat"), graindata, active, cells=grain_cells)
        p["read_in_graindata_from_file"] = "yes"
        p["grain_data_file"] = "grain.dat"
    if overrides:
        p.update(overrides)
    pfile = f"{name}.params"
    write_params(os.path.join(path, pfile), p)
    return Workload(path=path, params_file=pfile, params=p, dem=dem, dx=dx,
                    mannings=mannings, water_init=water_init,
                    rain=np.asarray(rain_mm_h))


# ---------------------------------------------------------------------------
# the five workloads (seeds 1000+k; ensemble members 4000+m)
# ---------------------------------------------------------------------------
def flow_only_256(path: str, size: int = 256, hours: int = 2, seed: int = 1001) -> Workload:
    """Plain method, flow only: size^2 DEM at 5 m, 0.01 planar slope toward
    the south, diamond-square noise (amplitude 5 m, roughness 0.6), a 3-cell
    channel carved 2 m deep, 10 mm/h rain for 1 h then 0, n=0.04,
    hflow_threshold 0.001 m, courant 0.7, south edge open only."""
    dem = planar_dem(size, size, 5.0, 0.01, 5.0, 0.6, seed)
    rain = np.array([10.0, 0.0])
    return make_workload(path, dem, 5.0, rain, 60, hours, name="flow256")


def triangular_storm(hours: float, peak: float, step_min: float) -> np.ndarray:
    n = int(round(hours * 60 / step_min))
    t = (np.arange(n) + 0.5) / n
    return peak * (1.0 - np.abs(2.0 * t - 1.0))


def dirichlet_props(nrows: int, ncols: int, rng: np.random.Generator,
                    alpha: float = 2.0) -> np.ndarray:
    return rng.dirichlet(np.full(9, alpha), size=(nrows, ncols))


def erosion_1024(path: str, size: int = 1024, hours: int = 24, seed: int = 1002) -> Workload:
    """Flow + erosion: size^2 at 10 m, noise amplitude 20 m, 9 grain fractions
    with seeded Dirichlet surface proportions, Wilcock-Crowe, 1 m active layer,
    6 h triangular storm peaking at 40 mm/h."""
    rng = np.random.Generator(np.random.PCG64(seed))
    dem = planar_dem(size, size, 10.0, 0.01, 20.0, 0.6, seed)
    props = dirichlet_props(size, size, rng)
    rain = triangular_storm(6, 40.0, 15)
    w = make_workload(path, dem, 10.0, rain, 15, hours, name="erode1024",
                      overrides={"hydro_model_only": "no",
                                 "active_layer_thickness": "1"},
                      graindata=props)
    w.extra["grain_d"] = EROSION_GRAIN_D
    return w


def erosion_arrays(size: int = 1024, wet_start: bool = False, hours: int = 24):
    """The flow + erosion workload: size^2 at 10 m, fractal relief 20 m on a
    1 % slope with a carved channel, 9 grain fractions (d = 0.0005 .. 0.128 m)
    in seeded Dirichlet surface proportions per cell, Wilcock-Crowe, 1 m active
    layer, a 6 h triangular storm peaking at 40 mm/h, `hours` simulated; from
    dry, or from a uniform 0.2 m (wet_start)."""
    dem = planar_dem(size, size, 10.0, 0.01, 20.0, 0.6, 1002)
    props = dirichlet_props(size, size, np.random.Generator(np.random.PCG64(1002)))
    arrays = {"dem": dem,
              "rain": np.vstack([triangular_storm(6, 40.0, 15)[:, None],
                                 np.zeros((4 * hours + 200, 1))]).astype(np.float32), "DX": 10.0,
              "graindata": graindata_array(props, 1.0)}
    over = {"hydro_model_only": "no", "active_layer_thickness": "1", "transport_law": "wilcock",
            "read_in_graindata_from_file": "yes", "grain_data_file": "grain.dat",
            "hc_grain_sizes": ",".join(repr(d) for d in EROSION_GRAIN_D)}
    if wet_start:
        arrays["water_init"] = np.where(dem > NODATA, 0.2, 0.0)
        over.update({"water_init_from_raster_on": "yes", "water_init_raster_file": "water_init.asc"})
    return arrays, params_text_dict(over, rain_step_min=15, run_hours=hours)


def erosion_run(path: str, size: int = 48, hours: int = 1, seed: int = 1002,
                failure_angle: float = 8.0, law: str = "wilcock",
                dem_noise: float = 0.0, grain_cover: float = 1.0,
                extra_params: Optional[Dict[str, str]] = None) -> Workload:
    """A whole flow + erosion run small enough to compare step by step: size^2
    at 10 m on a 2 % slope with 20 m of fractal relief and a 1 m channel, all
    four edges open, 9 grain fractions (d1 = 0.0005 m) in seeded Dirichlet
    surface proportions, a 1 m active layer, a uniform 0.1 m start, a
    triangular storm peaking at 40 mm/h over the run, and a low slope failure
    angle so that the hillslope processes move material (creep and global
    landsliding fire on the first iteration, in-channel landsliding every 10
    iterations).  The grain sizes go to the reference by its harness, to the
    oracle by extra['grain_d'] and to this build's driver by hc_grain_sizes.
    dem_noise adds that many metres of seeded Gaussian noise to the DEM (a
    perturbation for sensitivity floors).  grain_cover < 1 lists only that
    share of the cells (seeded) in the grain data file, in a seeded shuffled
    line order: the run then makes grain rows (addGS) as it goes.
    extra_params are added to the parameter file."""
    rng = np.random.Generator(np.random.PCG64(seed))
    dem = planar_dem(size, size, 10.0, 0.02, 20.0, 0.6, seed, channel_depth=1.0,
                     south_only=False)
    if dem_noise:
        noise = np.random.Generator(np.random.PCG64(seed + 77)).standard_normal(dem.shape)
        dem = np.where(dem > NODATA, dem + dem_noise * noise, dem)
    props = dirichlet_props(size, size, rng)
    cells = None
    if grain_cover < 1.0:
        crng = np.random.Generator(np.random.PCG64(seed + 55))
        cells = np.argwhere(crng.random((size, size)) < grain_cover)
        cells = cells[crng.permutation(len(cells))]
    over = {"hydro_model_only": "no", "active_layer_thickness": "1",
            "topmodel_m_value": "0.0005", "transport_law": law,
            "slope_failure_thresh": repr(failure_angle),
            "hc_grain_sizes": ",".join(repr(d) for d in EROSION_GRAIN_D)}
    over.update(extra_params or {})
    w = make_workload(path, dem, 10.0, triangular_storm(hours, 40.0, 15), 15, hours,
                      overrides=over, water_init=np.where(dem > NODATA, 0.1, 0.0),
                      graindata=props, grain_cells=cells, name="erun")
    w.extra["grain_d"] = EROSION_GRAIN_D
    return w


def complex_rain(path: str, rfnum: int = 1, size: int = 40, seed: int = 1021,
                 hours: int = 1) -> Workload:
    """spatially_complex_rainfall_on (a per-cell TOPMODEL): size^2 at 5 m,
    a 2 h triangular storm peaking at 30 mm/h in 10 min steps, 1 cm of water
    to start; with rfnum > 1 a hydroindex of rfnum bands of rows, band k's
    rain (1 + k/2) times the first's."""
    dem = planar_dem(size, size, 5.0, 0.01, 5.0, 0.6, seed)
    rain = triangular_storm(2, 30.0, 10)
    hydro = None
    if rfnum > 1:
        hydro = np.where(dem > NODATA, 1 + (np.arange(size)[:, None] * rfnum) // size, -9999)
        hydro = np.broadcast_to(hydro, dem.shape).astype(np.int32)
        rain = np.stack([rain * (1.0 + 0.5 * k) for k in range(rfnum)], axis=1)
    return make_workload(path, dem, 5.0, rain, 10, hours, hydroindex=hydro,
                         water_init=np.where(dem > NODATA, 0.01, 0.0),
                         overrides={"spatially_complex_rainfall_on": "yes"}, name="cplx")


def reach_inputs(path: str, erosion: bool = False, size: int = 40, seed: int = 1031,
                 hours: int = 1, div: int = 2, sed_scale: float = 1.0) -> Workload:
    """reach_mode with two input points on the north edge fed from reach
    files (16 columns: 1 = discharge m^3/s, 5..13 = sediment of fractions
    1..9, times ``sed_scale``), 15 min data steps, inputs divided by ``div``;
    1 cm of water to start; with ``erosion`` 9 grain fractions."""
    rng = np.random.Generator(np.random.PCG64(seed))
    dem = planar_dem(size, size, 5.0, 0.01, 5.0, 0.6, seed, south_only=False)
    steps = hours * 4 + 8
    t = np.arange(steps)
    files = {}
    for k, (x, y) in enumerate(((2, size // 3), (2, (2 * size) // 3))):
        rows = np.zeros((steps, 16))
        rows[:, 0] = t * 15
        rows[:, 1] = (0.4 + 0.2 * k) * np.sin(np.pi * np.minimum(t, 6) / 6.0) ** 2 + 0.02
        rows[:, 5:14] = rng.random((steps, 9)) * 0.002 * (1 + k) * sed_scale
        files[f"reach{k + 1}.txt"] = (rows, x, y)
    over = {"reach_mode": "yes", "divide_inputs_by": str(div), "reach_input_data_timestep": "15"}
    for k, (name, (rows, x, y)) in enumerate(files.items()):
        over.update({f"reach{k + 1}_input_on": "yes", f"reach{k + 1}_input_file": name,
                     f"reach{k + 1}_x": str(x), f"reach{k + 1}_y": str(y)})
    kw = {}
    if erosion:
        over.update({"hydro_model_only": "no", "active_layer_thickness": "0.3"})
        kw["graindata"] = dirichlet_props(size, size, rng)
    w = make_workload(path, dem, 5.0, np.array([0.0, 0.0]), 60, hours, overrides=over,
                      water_init=np.where(dem > NODATA, 0.01, 0.0), name="reach", **kw)
    for name, (rows, x, y) in files.items():
        np.savetxt(os.path.join(path, name), rows, fmt="%.9g")
    return w


def gamma_rain(hours: float, step_min: float, shape: float, mean: float,
               rng: np.random.Generator, ncells: int = 1) -> np.ndarray:
    n = int(round(hours * 60 / step_min))
    return rng.gamma(shape, mean / shape, size=(n, ncells))


def lognormal_mannings(nrows: int, ncols: int, median: float, sigma: float,
                       rng: np.random.Generator) -> np.ndarray:
    return median * np.exp(sigma * rng.standard_normal((nrows, ncols)))


def grid4096_dem(size: int = 4096, seed: int = 1003) -> np.ndarray:
    """The large single grid's DEM: 2 m cells, 0.2 % slope to the south edge,
    fractal relief of 30 m (roughness 0.6) and a D8-consistent drainage
    network carved 3 m deep (carve_d8_network on 64-cell blocks)."""
    dem = planar_dem(size, size, 2.0, 0.002, 30.0, 0.6, seed, channel_depth=0.0)
    block = max(8, min(64, size // 16))
    return carve_d8_network(dem, block=block, depth=3.0)[0]


def large_4096(path: str, size: int = 4096, hours: int = 12, seed: int = 1003,
               wet_depth: Optional[float] = None) -> Workload:
    """Large single grid: size^2 at 2 m, fractal amplitude 30 m with carved
    drainage, log-normal Manning n (median 0.035, sigma 0.3), 12 h of
    Gamma(0.8, mean 8 mm/h) 15-minute rain.  ``wet_depth`` adds a uniform
    initial depth (fully wet measurement state)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    dem = grid4096_dem(size, seed)
    n = lognormal_mannings(size, size, 0.035, 0.3, rng)
    rain = gamma_rain(hours, 15, 0.8, 8.0, rng)
    wi = None
    if wet_depth is not None:
        wi = np.full((size, size), wet_depth)
        wi[dem == NODATA] = 0.0
    return make_workload(path, dem, 2.0, rain, 15, hours, mannings=n,
                         water_init=wi, name="grid4096")


def params_text_dict(overrides: Dict[str, str] = None, rain_step_min: float = 60,
                     run_hours: int = 1) -> Dict[str, str]:
    """Parameter dictionary for in-memory runs (file keys are placeholders)."""
    p = dict(BASE_PARAMS)
    p.update({"read_path": "./", "write_path": "./", "read_fname": "dem",
              "write_fname": "run.dat", "rainfall_data_file": "rain.txt",
              "rain_data_time_step": "%g" % rain_step_min,
              "max_run_duration": str(int(run_hours))})
    if overrides:
        p.update(overrides)
    return p


def grid4096_arrays(size: int = 4096, seed: int = 1003, wet_depth: Optional[float] = 0.2,
                    hours: int = 12):
    """The large single-grid workload in memory: size^2 cells at 2 m, fractal
    relief (30 m) on a 0.2 % slope to the south edge with a carved D8 network,
    log-normal Manning n (median 0.035, sigma 0.3), 12 h of 15-minute
    Gamma(0.8, mean 8 mm/h) rain, flow only.  With ``wet_depth`` every cell
    starts at that uniform depth (fully wet measurement state); with None the
    grid starts dry.  Input points use the reference's default
    min_q_for_depth_calc (0.01)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    dem = grid4096_dem(size, seed)
    n = lognormal_mannings(size, size, 0.035, 0.3, rng)
    rain = gamma_rain(hours, 15, 0.8, 8.0, rng)
    rain = np.vstack([rain, np.zeros((200, 1))]).astype(np.float32)
    over = {"spatially_variable_mannings_on": "yes", "spatial_mannings_dem_file": "mannings.asc",
            "min_q_for_depth_calc": "0.01"}
    arrays = {"dem": dem, "mannings": n, "rain": rain, "DX": 2.0}
    if wet_depth is not None:
        arrays["water_init"] = np.where(dem > NODATA, wet_depth, 0.0)
        over.update({"water_init_from_raster_on": "yes",
                     "water_init_raster_file": "water_init.asc"})
    params = params_text_dict(over, rain_step_min=15, run_hours=hours)
    return arrays, params


def ensemble_members(n: int = 64, lo: int = 128, hi: int = 1024, hours: int = 6,
                     wet_depth: Optional[float] = None, seed0: int = 4000,
                     sizes: Optional[List[int]] = None):
    """The batched-ensemble workload in memory: n independent catchments with
    side lengths L = round(exp(U(ln lo, ln hi))) cells at 5 m (member m seeded
    seed0+m), each a fractal DEM (5 m relief) on a 1 % slope to the south edge
    with a carved channel, a uniform Manning n ~ U(0.02, 0.08), and `hours` of
    15-minute Gamma(0.8, mean 8 mm/h) rain.  With wet_depth every member starts
    at that uniform depth (fully wet measurement state); otherwise dry.
    Returns [(arrays, params)]."""
    out = []
    for m in range(n):
        rng = np.random.Generator(np.random.PCG64(seed0 + m))
        L = sizes[m] if sizes is not None else int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
        dem = planar_dem(L, L, 5.0, 0.01, 5.0, 0.6, seed0 + m, channel_depth=2.0)
        nman = float(rng.uniform(0.02, 0.08))
        rain = gamma_rain(hours, 15, 0.8, 8.0, rng)
        rain = np.vstack([rain, np.zeros((200, 1))]).astype(np.float32)
        over = {"mannings_n": repr(nman)}
        arrays = {"dem": dem, "rain": rain, "DX": 5.0}
        if wet_depth is not None:
            arrays["water_init"] = np.where(dem > NODATA, wet_depth, 0.0)
            over.update({"water_init_from_raster_on": "yes",
                         "water_init_raster_file": "water_init.asc"})
        out.append((arrays, params_text_dict(over, rain_step_min=15, run_hours=hours)))
    return out


def pad(a: np.ndarray, fill: float) -> np.ndarray:
    """An interior (imax x jmax) array on the reference's padded layout
    ((imax+2) x (jmax+2), a one-cell ring of `fill`)."""
    out = np.full((a.shape[0] + 2, a.shape[1] + 2), fill, dtype=a.dtype)
    out[1:-1, 1:-1] = a
    return out


def smooth_noise(rng: np.random.Generator, shape, sigma: float) -> np.ndarray:
    """White noise convolved with a Gaussian of `sigma` cells (periodic FFT
    convolution), scaled to unit max."""
    xi = rng.standard_normal(shape)
    ky = np.fft.fftfreq(shape[0])[:, None]
    kx = np.fft.fftfreq(shape[1])[None, :]
    g = np.exp(-2 * (np.pi * sigma) ** 2 * (kx ** 2 + ky ** 2))
    s = np.real(np.fft.ifft2(np.fft.fft2(xi) * g))
    return s / np.abs(s).max()


def adjoint_case(size: int = 512, steps: int = 2000, seed: int = 1005, dx: float = 5.0,
                 channel: float = 0.0,
                 relief: float = 5.0, slope: float = 0.01, depth: float = 0.1,
                 median_n: float = 0.035, sigma_n: float = 0.3, sigma_pert: float = 8.0):
    """The adjoint workload: size^2 cells at dx, a fractal DEM (`relief` m,
    roughness 0.6) on a `slope` to the south edge (with `channel` > 0, a
    channel carved that deep down the middle column to the south edge),
    log-normal Manning n, a uniform `depth` m start, 8 rain blocks (2 rows x
    4 columns of the grid) under a rain rate of (10 + 30 sin^2(t/300)) mm/h
    per step t, and the target's n field: n (1 + 0.1 s), s Gaussian-smoothed
    noise (sigma_pert cells) scaled to unit max.  Returns a dict of interior
    arrays (dem, n, n_target, water), the padded block map (-1 off the grid)
    and the per-step rain rates in m/s."""
    rng = np.random.Generator(np.random.PCG64(seed))
    dem = planar_dem(size, size, dx, slope, relief, 0.6, seed, channel_depth=channel)
    # n is held at FP32 precision (the gradient is with respect to FP32 n)
    n = lognormal_mannings(size, size, median_n, sigma_n, rng).astype(np.float32).astype(np.float64)
    npert = n * (1 + 0.1 * smooth_noise(rng, (size, size), sigma_pert))
    block = -np.ones((size + 2, size + 2), np.int32)
    r = (np.arange(size) * 2) // size
    c = (np.arange(size) * 4) // size
    block[1:-1, 1:-1] = r[:, None] * 4 + c[None, :]
    t = np.arange(steps)
    rain = (10.0 + 30.0 * np.sin(t / 300.0) ** 2) / 3.6e6
    water = np.where(dem > NODATA, depth, 0.0)
    return {"dem": dem, "n": n, "n_target": npert, "water": water, "block": block,
            "rain": rain, "DX": dx}

#!/usr/bin/env python3
"""Test infrastructure: the CPU baseline of the large single grid, timed here.

Times whole loop bodies (src/main.cpp:139-204) on the 4096^2 fully wet state
(inputs.grid4096_arrays: 2 m cells, fractal DEM with a carved D8 network,
log-normal Manning n, 15-minute Gamma rain, a uniform 0.2 m start, every cell
active) for:
  reference_o3  the reference compiled from its own sources with
                -O3 -march=native -fopenmp (oracle/Makefile ref-timing)
  reference_g   the same sources with the shipped Makefile's flags (-g, no
                optimisation)
  port          the oracle (oracle/liboracle.so, the FP64 OpenMP restatement
                that bench.py times on the GPU box's host cores)
each in its own process with OMP_NUM_THREADS threads and OMP_PROC_BIND=close,
after 6 warm-up bodies (the active mask is first built at counter 5).  The
reference reads the workload from files written once to a scratch directory
(its own ASCII formats, elevations at %.17g so every run reads the same FP64
values).  Writes profiles/cpu_baseline_4096.json with the port's speed ratio
to the optimised reference.

With --erosion: the flow + erosion workload of the benchmark's erosion
section at 256^2 (10 m cells, 20 m fractal relief on a 1 % slope, 9 grain
fractions in seeded Dirichlet proportions in every cell's grain row,
Wilcock-Crowe, 1 m active layer, the 40 mm/h storm, a uniform 0.2 m start),
whole loop bodies after 3 warm-up bodies (the first runs creep and global
landsliding); writes profiles/cpu_baseline_erosion_256.json.

Needs the reference tree (runs only where oracle/_ref was built).
Usage: python oracle/time_cpu.py [--threads 8] [--bodies 20] [--workdir DIR] [--erosion]
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
OUT = os.path.join(ROOT, "profiles", "cpu_baseline_4096.json")
OUT_EROSION = os.path.join(ROOT, "profiles", "cpu_baseline_erosion_256.json")
EROSION_SIZE = 256


def _package():
    pkg = os.path.join(ROOT, "hail-caesar_amd")
    spec = importlib.util.spec_from_file_location(
        "hail_caesar_amd", os.path.join(pkg, "__init__.py"), submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["hail_caesar_amd"] = mod
    spec.loader.exec_module(mod)
    import hail_caesar_amd.inputs as inputs
    return inputs


def write_workload(d: str) -> None:
    inputs = _package()
    arrays, params = inputs.grid4096_arrays()
    os.makedirs(os.path.join(d, "out"), exist_ok=True)
    dx = arrays["DX"]
    inputs.write_asc(os.path.join(d, "dem.asc"), arrays["dem"], dx)
    inputs.write_asc(os.path.join(d, "mannings.asc"), arrays["mannings"], dx)
    inputs.write_asc(os.path.join(d, "water_init.asc"), arrays["water_init"], dx)
    inputs.write_rain(os.path.join(d, "rain.txt"), arrays["rain"][:-200], pad_rows=200)
    p = dict(params)
    p["write_path"] = "./out/"
    inputs.write_params(os.path.join(d, "grid4096.params"), p)
    with open(os.path.join(d, "done"), "w") as f:
        f.write("ok\n")


def erosion_arrays(size: int = EROSION_SIZE):
    """The benchmark's erosion workload (bench.py run_erosion) at `size`."""
    import numpy as np
    inputs = _package()
    dem = inputs.planar_dem(size, size, 10.0, 0.01, 20.0, 0.6, 1002)
    props = inputs.dirichlet_props(size, size, np.random.Generator(np.random.PCG64(1002)))
    arrays = {"dem": dem, "water_init": np.where(dem > inputs.NODATA, 0.2, 0.0),
              "rain": np.vstack([inputs.triangular_storm(6, 40.0, 15)[:, None],
                                 np.zeros((200, 1))]).astype(np.float32), "DX": 10.0,
              "props": props}
    params = inputs.params_text_dict({
        "hydro_model_only": "no", "active_layer_thickness": "1", "transport_law": "wilcock",
        "water_init_from_raster_on": "yes", "water_init_raster_file": "water_init.asc",
        "read_in_graindata_from_file": "yes", "grain_data_file": "grain.dat",
        "hc_grain_sizes": ",".join(repr(d) for d in inputs.EROSION_GRAIN_D)},
        rain_step_min=15, run_hours=24)
    return arrays, params


def write_erosion_workload(d: str) -> None:
    inputs = _package()
    arrays, params = erosion_arrays()
    os.makedirs(os.path.join(d, "out"), exist_ok=True)
    dx = arrays["DX"]
    inputs.write_asc(os.path.join(d, "dem.asc"), arrays["dem"], dx)
    inputs.write_asc(os.path.join(d, "water_init.asc"), arrays["water_init"], dx)
    inputs.write_rain(os.path.join(d, "rain.txt"), arrays["rain"][:-200], pad_rows=200)
    inputs.write_graindata(os.path.join(d, "grain.dat"), arrays["props"], 1.0)
    p = dict(params)
    p["write_path"] = "./out/"
    p["read_fname"] = "dem"
    inputs.write_params(os.path.join(d, "erosion.params"), p)
    with open(os.path.join(d, "done"), "w") as f:
        f.write("ok\n")


def child_erosion(kind: str, d: str, threads: int, bodies: int) -> dict:
    sys.path.insert(0, HERE)
    import numpy as np
    import pyoracle
    inputs = _package()
    t0 = time.perf_counter()
    if kind.startswith("reference"):
        pyoracle.REF_SO = os.path.join(HERE, "_ref", "libhcref_o3.so" if kind == "reference_o3"
                                       else "libhcref_g.so")
        m = pyoracle.RefModel(d, "erosion.params", threads=threads)
        for k, v in enumerate(inputs.EROSION_GRAIN_D):  # the reference's own keys lack d1..d9
            m.scalar(f"d{k + 1}", v)
    else:
        arrays, params = erosion_arrays()
        dem = arrays["dem"]
        p = pyoracle.params_from_dict(params, dem.shape[0], dem.shape[1], arrays["DX"],
                                      nthreads=threads, grain_d=inputs.EROSION_GRAIN_D)
        m = pyoracle.OracleModel(p)
        m.field("elev")[...] = pyoracle.pad(dem, -9999.0)
        m.field("water_depth")[1:-1, 1:-1] = arrays["water_init"]
        m.set_rain(arrays["rain"].astype("float64"))
        m.load_grain_arrays(inputs.graindata_array(arrays["props"], 1.0))
        m.L.orc_initialise(m.h)
    setup = time.perf_counter() - t0
    m.iterate(3)
    t = time.perf_counter()
    m.iterate(bodies)
    el = time.perf_counter() - t
    cells = EROSION_SIZE * EROSION_SIZE
    return {"kind": kind, "threads": threads, "bodies": bodies, "seconds": el,
            "cell_steps_per_s": cells * bodies / el, "setup_s": setup,
            "elev_checksum": float(np.asarray(m.field("elev"))[1:-1, 1:-1].sum())}


def child(kind: str, d: str, threads: int, bodies: int) -> dict:
    sys.path.insert(0, HERE)
    import pyoracle
    t0 = time.perf_counter()
    if kind.startswith("reference"):
        pyoracle.REF_SO = os.path.join(HERE, "_ref", "libhcref_o3.so" if kind == "reference_o3"
                                       else "libhcref_g.so")
        m = pyoracle.RefModel(d, "grid4096.params", threads=threads)
    else:
        inputs = _package()
        arrays, params = inputs.grid4096_arrays()
        dem = arrays["dem"]
        p = pyoracle.params_from_dict(params, dem.shape[0], dem.shape[1], arrays["DX"], nthreads=threads)
        m = pyoracle.OracleModel(p)
        m.field("elev")[...] = pyoracle.pad(dem, -9999.0)
        m.field("mannings")[1:-1, 1:-1] = arrays["mannings"]
        m.field("water_depth")[1:-1, 1:-1] = arrays["water_init"]
        m.set_rain(arrays["rain"].astype("float64"))
        m.L.orc_initialise(m.h)
    setup = time.perf_counter() - t0
    m.iterate(6)
    t = time.perf_counter()
    m.iterate(bodies)
    el = time.perf_counter() - t
    cells = 4096 * 4096
    return {"kind": kind, "threads": threads, "bodies": bodies, "seconds": el,
            "cell_steps_per_s": cells * bodies / el, "setup_s": setup}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--bodies", type=int, default=20)
    ap.add_argument("--workdir", default="/tmp/hc_cpu4096")
    ap.add_argument("--child", default=None)
    ap.add_argument("--erosion", action="store_true")
    args = ap.parse_args()
    if args.erosion and args.workdir == "/tmp/hc_cpu4096":
        args.workdir = "/tmp/hc_cpu_erosion256"
    if args.child:
        f = child_erosion if args.erosion else child
        print(json.dumps(f(args.child, args.workdir, args.threads, args.bodies)), flush=True)
        return
    if args.erosion:
        return main_erosion(args)
    if not os.path.exists(os.path.join(args.workdir, "done")):
        print("writing the 4096^2 workload files ...", file=sys.stderr, flush=True)
        write_workload(args.workdir)
    env = dict(os.environ, OMP_NUM_THREADS=str(args.threads), OMP_PROC_BIND="close")
    res = {}
    for kind, bodies in (("reference_o3", args.bodies), ("port", args.bodies),
                         ("reference_g", max(3, args.bodies // 5))):
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--child", kind,
                            "--threads", str(args.threads), "--bodies", str(bodies),
                            "--workdir", args.workdir], env=env, capture_output=True, text=True)
        if r.returncode != 0:
            sys.exit(f"{kind} failed:\n{r.stderr[-2000:]}")
        res[kind] = json.loads(r.stdout.strip().splitlines()[-1])
        print(json.dumps(res[kind]), file=sys.stderr, flush=True)
    cpu = ""
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                cpu = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    out = {"workload": "4096^2 at 2 m, fractal DEM + D8-carved network, log-normal n, 15-min "
                       "Gamma rain, fully wet (0.2 m); whole loop bodies after 6 warm-up bodies",
           "host": {"cpu": cpu, "threads": args.threads, "omp_proc_bind": "close"},
           "runs": res,
           "port_over_reference": res["port"]["cell_steps_per_s"] / res["reference_o3"]["cell_steps_per_s"],
           "reference_o3_over_g": res["reference_o3"]["cell_steps_per_s"] / res["reference_g"]["cell_steps_per_s"]}
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(json.dumps(out, indent=1))


def main_erosion(args):
    if not os.path.exists(os.path.join(args.workdir, "done")):
        print("writing the 256^2 erosion workload files ...", file=sys.stderr, flush=True)
        write_erosion_workload(args.workdir)
    env = dict(os.environ, OMP_NUM_THREADS=str(args.threads), OMP_PROC_BIND="close")
    res = {}
    for kind in ("reference_o3", "port"):
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--erosion", "--child", kind,
                            "--threads", str(args.threads), "--bodies", str(args.bodies),
                            "--workdir", args.workdir], env=env, capture_output=True, text=True)
        if r.returncode != 0:
            sys.exit(f"{kind} failed:\n{r.stderr[-2000:]}")
        res[kind] = json.loads(r.stdout.strip().splitlines()[-1])
        print(json.dumps(res[kind]), file=sys.stderr, flush=True)
    cpu = ""
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                cpu = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    out = {"workload": f"{EROSION_SIZE}^2 flow + erosion run (the benchmark's erosion workload at "
                       f"{EROSION_SIZE}^2); whole loop bodies after 3 warm-up bodies",
           "host": {"cpu": cpu, "threads": args.threads, "omp_proc_bind": "close"},
           "runs": res,
           "port_over_reference": res["port"]["cell_steps_per_s"] / res["reference_o3"]["cell_steps_per_s"]}
    with open(OUT_EROSION, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()

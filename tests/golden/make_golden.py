#!/usr/bin/env python3
"""Write the golden vectors under tests/golden/ from the compiled reference.

Runs only where oracle/_ref/libhcref.so was built from the reference sources
(oracle/Makefile `make ref`).  Every fixture is input + expected output data:
  flow_step_64.npz            a wet 64^2 flow-only state after 400 reference
                              iterations, then one flow_route, depth_update,
                              scan_area and water_flux_out call
  flow_step_spatial_n_48.npz  the same on 48^2 with per-cell Manning n
  flow_run_64.npz             a whole 2 h 64^2 flow-only run: per-step
                              outlet discharge, clock and flow timestep, final
                              h/qx/qy
  config1_run_256.npz         the same for the 256^2 configuration (final
                              fields subsampled, with whole-field sums)
  grainz_run_48.npz           a whole 48^2 erosion run whose grain data file
                              lists half the cells in shuffled order: the
                              final grain-row numbers (index) and
                              grain_array_tot, and of every Grainz file the
                              reference wrote (LGM.cpp:59-118) its name, its
                              x y index columns and its first 40 lines
The reference runs single-threaded (its OpenMP loops race on shared members
with spatial n and in erode).

Usage: python tests/golden/make_golden.py
"""
from __future__ import annotations

import importlib.util
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def load_package():
    pkg = os.path.join(ROOT, "hail-caesar_amd")
    spec = importlib.util.spec_from_file_location(
        "hail_caesar_amd", os.path.join(pkg, "__init__.py"), submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["hail_caesar_amd"] = mod
    spec.loader.exec_module(mod)


load_package()
import hail_caesar_amd.inputs as inputs  # noqa: E402
from pyoracle import RefModel  # noqa: E402


def per_call(w, steps, name, extra=None):
    r = RefModel(w.path, w.params_file, threads=1)
    r.iterate(steps)
    out = {
        "params_json": json.dumps(w.params), "imax": w.dem.shape[0], "jmax": w.dem.shape[1],
        "DX": w.dx, "elev": r.field("elev").copy(), "h0": r.field("water_depth").copy(),
        "qx0": r.field("qx").copy(), "qy0": r.field("qy").copy(), "mask0": r.mask(),
        "maxdepth0": r.scalar("maxdepth"), "time_step0": r.scalar("time_step"),
    }
    if w.params.get("spatially_variable_mannings_on") == "yes":
        out["n"] = r.field("spat_var_mannings").copy()
    out["dt"] = r.call("get_flow_timestep")
    r.call("flow_route")
    out["qx1"] = r.field("qx").copy()
    out["qy1"] = r.field("qy").copy()
    out["maxdepth1"] = r.call("depth_update")
    out["h1"] = r.field("water_depth").copy()
    r.call("scan_area")
    out["mask1"] = r.mask()
    out["temptot1"] = r.call("water_flux_out")
    out["h2"] = r.field("water_depth").copy()
    np.savez_compressed(os.path.join(HERE, name), **out)
    print("wrote", name)


def full_run(w, name):
    r = RefModel(w.path, w.params_file, threads=1)
    tt, cy, dt = r.iterate(10 ** 7)
    rain = np.loadtxt(os.path.join(w.path, w.params["rainfall_data_file"]), ndmin=2)
    np.savez_compressed(os.path.join(HERE, name), params_json=json.dumps(w.params), dem=w.dem,
                        DX=w.dx, rain=rain, temptot=tt, cycle=cy, dtflow=dt,
                        water_depth=r.field("water_depth").copy(), qx=r.field("qx").copy(),
                        qy=r.field("qy").copy())
    print("wrote", name, len(tt), "steps")


def config1_run(w, name, stride=4):
    """The plain flow-only configuration at its stated size (256^2, 2 h):
    per-step outlet discharge, clock and flow timestep over the whole run,
    and the final h/qx/qy on every `stride`-th row and column of the padded
    layout with whole-field sums (the full fields would be 1.5 MB).  The DEM
    is regenerated from its seed by the tests; its SHA-256 is kept to check
    that.  Uniform n and no erosion, so the reference's OpenMP loops have no
    race and it runs on 8 threads."""
    import hashlib
    r = RefModel(w.path, w.params_file, threads=8)
    tt, cy, dt = r.iterate(10 ** 7)
    out = {"params_json": json.dumps(w.params), "DX": w.dx, "stride": stride,
           "dem_sha256": hashlib.sha256(np.ascontiguousarray(w.dem).tobytes()).hexdigest(),
           "temptot": tt, "cycle": cy, "dtflow": dt}
    for f in ("water_depth", "qx", "qy"):
        a = r.field(f).copy()
        out[f + "_sub"] = a[::stride, ::stride]
        out[f + "_sum"] = float(a.sum())
        out[f + "_abssum"] = float(np.abs(a).sum())
        out[f + "_max"] = float(a.max())
    np.savez_compressed(os.path.join(HERE, name), **out)
    print("wrote", name, len(tt), "steps")


GRAINZ_PARAMS = {"write_grainsize_file": "yes", "grainsize_file": "grainz",
                 "raster_output_interval": "30", "dem_write_extension": "asc"}


def grainz_run(name):
    """The erosion-run workload with half the cells in its grain data file
    (inputs.erosion_run(grain_cover=0.5)) and Grainz output every 30
    simulated minutes, run to its end."""
    with tempfile.TemporaryDirectory() as tmp:
        w = inputs.erosion_run(os.path.join(tmp, "gz"), grain_cover=0.5, extra_params=GRAINZ_PARAMS)
        r = RefModel(w.path, w.params_file, threads=1)
        r.scalar("d1", inputs.EROSION_GRAIN_D[0])
        tt, _, _ = r.iterate(10 ** 7)
        out = {"index": r.field_int("index").copy(),
               "grain_array_tot": r.scalar("grain_array_tot"), "steps": len(tt)}
        odir = os.path.join(w.path, "out")
        files = sorted(f for f in os.listdir(odir) if f.startswith("grainz"))
        out["files"] = np.array(files)
        for k, f in enumerate(files):
            with open(os.path.join(odir, f)) as fh:
                lines = fh.read().split("\n")
            assert lines[-1] == ""
            lines = lines[:-1]
            out[f"cols{k}"] = np.array([[int(t) for t in ln.split(" ")[:3]] for ln in lines],
                                       dtype=np.int32).reshape(-1, 3)
            out[f"head{k}"] = np.array(lines[:40])
        np.savez_compressed(os.path.join(HERE, name), **out)
        print("wrote", name, len(tt), "steps,", files)


def erode_case(name, n, steps, law, ncalls=2):
    """A wet, steep grain-bearing state (all four edges open) after `steps`
    reference iterations in erosion mode, then `ncalls` erode() calls."""
    with tempfile.TemporaryDirectory() as tmp:
        rng = np.random.Generator(np.random.PCG64(77))
        dem = inputs.planar_dem(n, n, 10.0, 0.02, 6.0, 0.6, 1002, channel_depth=1.0, south_only=False)
        props = inputs.dirichlet_props(n, n, rng)
        wi = np.where(dem > -9999, 0.15, 0.0)
        w = inputs.make_workload(os.path.join(tmp, "e"), dem, 10.0, inputs.triangular_storm(1, 40.0, 15),
                                 15, 1, overrides={"hydro_model_only": "no", "active_layer_thickness": "1",
                                                   "topmodel_m_value": "0.0005", "transport_law": law},
                                 water_init=wi, graindata=props, name="er")
        r = RefModel(w.path, w.params_file, threads=1)
        r.scalar("d1", inputs.EROSION_GRAIN_D[0])
        r.iterate(steps)
        before = r.snapshot()
        out = {"params_json": json.dumps(w.params), "imax": n, "jmax": n, "DX": 10.0,
               "grain_d": np.array(inputs.EROSION_GRAIN_D)}
        out.update({"in_" + k: v for k, v in before.items()})
        mult = float(before["erode_mult"])
        for c in range(ncalls):
            out[f"tempbmax{c}"] = r.call("erode", mult)
            out[f"time_step{c}"] = r.scalar("time_step")
        after = r.snapshot()
        for k in ("elev", "Vsusptot", "Tau", "Vel", "ss", "bedrock", "grain", "strata", "hasgrain",
                  "sr", "sl", "su", "sd", "globalsediq", "sum_grain", "time_step"):
            out["out_" + k] = after[k]
        out["mult"] = mult
        np.savez_compressed(os.path.join(HERE, name), **out)
        print("wrote", name)


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "grainz":
        grainz_run("grainz_run_48.npz")
        return
    if len(sys.argv) > 1 and sys.argv[1] == "config1":
        with tempfile.TemporaryDirectory() as tmp:
            config1_run(inputs.flow_only_256(os.path.join(tmp, "d"), size=256, hours=2),
                        "config1_run_256.npz")
        return
    if not RefModel.available():
        sys.exit("compiled reference not built (make -C oracle ref)")
    with tempfile.TemporaryDirectory() as tmp:
        w = inputs.flow_only_256(os.path.join(tmp, "a"), size=64, hours=2)
        per_call(w, 400, "flow_step_64.npz")
        rng = np.random.Generator(np.random.PCG64(5))
        dem = inputs.planar_dem(48, 48, 5.0, 0.01, 5.0, 0.6, 1001)
        n = inputs.lognormal_mannings(48, 48, 0.035, 0.3, rng)
        w2 = inputs.make_workload(os.path.join(tmp, "b"), dem, 5.0, np.array([20.0, 5.0, 0.0]), 30, 1,
                                  mannings=n, water_init=np.where(dem > -9999, 0.02, 0.0), name="spn")
        per_call(w2, 300, "flow_step_spatial_n_48.npz")
        w3 = inputs.flow_only_256(os.path.join(tmp, "c"), size=64, hours=2)
        full_run(w3, "flow_run_64.npz")
        w4 = inputs.flow_only_256(os.path.join(tmp, "d"), size=256, hours=2)
        config1_run(w4, "config1_run_256.npz")
    erode_case("erode_wilcock_32.npz", 32, 60, "wilcock")
    erode_case("erode_einstein_24.npz", 24, 40, "einstein")
    grainz_run("grainz_run_48.npz")


if __name__ == "__main__":
    main()

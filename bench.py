#!/usr/bin/env python3
"""Benchmark of the per-timestep hot path on one MI355X.

Workload: the large single grid (4096^2 cells at 2 m, fractal relief,
log-normal Manning n, Gamma rainfall, flow only) from a fully wet state (a
uniform 0.2 m initial depth).  One step is one iteration of the reference's
loop body (src/main.cpp:139-204): Courant timestep, clocks, rainfall input,
flow_route + depth_update (one fused kernel), scan_area every 5th step and
water_flux_out, all device-resident and replayed from a captured HIP graph.

The headline value is FP32-storage cell-steps/s; the FP64 plain form runs on
the same inputs and is reported beside it, with the FP32/FP64 ratio.  The
dominant kernel's duration is measured with HIP events on the context stream
in a second pass of the same steps.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "hail-caesar_amd")
HBM_PEAK_GBS = 8000.0
BASELINE = os.path.join(ROOT, "BASELINE.json")
TRAFFIC = os.path.join(ROOT, "profiles", "hbm_traffic.json")


def load_package():
    if "hail_caesar_amd" in sys.modules:
        return sys.modules["hail_caesar_amd"]
    spec = importlib.util.spec_from_file_location(
        "hail_caesar_amd", os.path.join(PKG, "__init__.py"), submodule_search_locations=[PKG])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["hail_caesar_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def bytes_per_cell_step(precision: str, spatial_n: bool = True) -> int:
    """Algorithmic bytes of one fused flux+depth step per cell: reads h, qx,
    qy, z (and n), writes h, qx, qy, each field 8 bytes in the FP64 form and 4
    in the FP32 form (whose elevations are 32-bit fixed point)."""
    esz = 4 if precision == "fp32" else 8
    return esz * (3 + 3 + 1 + (1 if spatial_n else 0))


def copy_probe(mib: int):
    """A plain device copy of known size (torch's elementwise copy kernel),
    run under the same hardware counters as the fused kernel so that their
    HBM byte counts can be scaled to true bytes."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": 4 * n, "copy_probe_bytes_written": 4 * n}),
          file=sys.stderr, flush=True)
    del a, b


def run_gpu(precision: str, arrays, params, steps: int, warmup: int, use_graph: bool = True,
            tiles=None):
    from hail_caesar_amd import hc
    from hail_caesar_amd.model import Model
    prec = hc.HC_FP32 if precision == "fp32" else hc.HC_FP64
    t0 = time.time()
    m = Model(None, None, precision=prec, arrays=arrays, params=params, device_loop=True,
              use_graph=use_graph, elev_fixed=precision == "fp32")
    counter0 = m.scalar("counter")
    setup_s = time.time() - t0
    m.iterate(warmup)
    # the project's tile tuner picks rows per wavefront on the warmed state
    # (isolated launches on a fixed state); the pick and the usual
    # candidates are then timed inside the loop itself, where each step reads
    # what the step before wrote, and the fastest there is kept (untimed steps)
    ctx = m.context()
    if tiles:  # fixed tiling (profiler runs that must all see the same kernel)
        rows, strip, tune_us = tiles[0], tiles[1], None
        cands = [tuple(tiles)]
    else:
        rows, strip, tune_us = ctx.tune_tiles(0.05)
        cands = [(rows, strip), (4, 2), (8, 1), (4, 4), (8, 2), (16, 1)]
    loop_us = {}
    m.set_graph(False)
    for rs in dict.fromkeys(cands):
        ctx.set_tiles(*rs)
        ctx.loop_profile(True)
        m.iterate(20)
        t, n = ctx.loop_kernel_time()
        ctx.loop_profile(False)
        loop_us[rs] = 1e3 * t / max(n, 1)
    m.set_graph(use_graph)
    rows, strip = min(loop_us, key=loop_us.get)
    ctx.set_tiles(rows, strip)
    m.iterate(2)
    # timed region: K loop bodies, synchronised on both sides
    m.synchronize()
    t1 = time.perf_counter()
    m.iterate(steps)
    m.synchronize()
    wall = time.perf_counter() - t1
    # kernel timing pass: the same loop without graphs, HIP events around
    # every fused flow_route+depth_update launch on the context stream
    ctx.loop_profile(True)
    m.set_graph(False)
    m.iterate(steps)
    ktot, kn = ctx.loop_kernel_time()
    ctx.loop_profile(False)
    m.set_graph(use_graph)
    mask = m.download("mask")
    active = float(mask[1:-1, 1:-1].mean())
    h = m.download("water_depth")
    return {
        "wall_s": wall, "setup_s": setup_s, "kernel_ms": ktot / max(kn, 1), "kernel_launches": kn,
        "active_fraction": active, "h": h, "counter": m.scalar("counter"), "cycle": m.scalar("cycle"),
        "bodies": int(round(m.scalar("counter") - counter0)),
        "model": m, "tile_rows": rows, "tile_strip": strip, "tune_us": tune_us,
        "loop_tile_us": {f"{r}x{s}": round(u, 1) for (r, s), u in loop_us.items()},
    }


def _field_errors(a, b):
    """Interior-cell errors of field a against reference b: max-norm relative
    to max|b|, L2 relative, and the share of cells off by more than 1e-4 of
    max|b|."""
    a, b = a[1:-1, 1:-1], b[1:-1, 1:-1]
    d = np.abs(a - b)
    scale = float(np.abs(b).max()) or 1.0
    return {"max_rel": float(d.max() / scale),
            "l2_rel": float(np.sqrt((d ** 2).sum() / max((b ** 2).sum(), 1e-300))),
            "frac_cells_over_1e-4": float((d > 1e-4 * scale).mean()),
            "p99_9_abs": float(np.percentile(d, 99.9))}


def fp32_error(r32, r64, arrays, params):
    """Per-field error of the FP32 form against the FP64 form after the same
    loop bodies, beside the FP64 form's own sensitivity floor: the FP64 form
    started from the DEM plus 1e-8 m of noise (an FP64-scale perturbation).
    The depth limiter (outflow > h/4 resets the flux to h/5) is
    discontinuous, so any perturbation flips it on a few cells near the
    threshold; the floor shows how far that alone carries.  meets_max_1e-4
    states whether the depth's max-norm error is within 1e-4 of the largest
    depth; within_fp64_floor whether every depth figure is within the
    floor's.  A second floor starts the FP64 form from its inputs rounded to
    float32 (DEM, initial depth, Manning n: the FP32 form stores n in FP32),
    a sub-ulp perturbation of the FP32 grid; within_fp64_floor_f32_inputs
    compares against it."""
    from hail_caesar_amd import hc
    from hail_caesar_amd.model import Model
    # both forms to the same number of loop bodies (their tile trials differ)
    m32, m64 = r32["model"], r64["model"]
    nsteps = max(r32["bodies"], r64["bodies"])
    for r in (r32, r64):
        if r["bodies"] < nsteps:
            r["model"].iterate(nsteps - r["bodies"])
    pert = dict(arrays)
    dem = arrays["dem"]
    noise = 1e-8 * np.random.Generator(np.random.PCG64(3)).standard_normal(dem.shape)
    pert["dem"] = dem + np.where(dem > -9999.0, noise, 0.0)
    mf = Model(None, None, precision=hc.HC_FP64, arrays=pert, params=params, device_loop=True,
               use_graph=True)
    mf.iterate(nsteps)
    mf.synchronize()
    out = {"steps": nsteps, "sim_minutes_fp64": m64.scalar("cycle"),
           "sim_minutes_fp32": m32.scalar("cycle"), "fields": {}, "fp64_floor": {}}
    for f in ("water_depth", "qx", "qy"):
        ref = m64.download(f)
        out["fields"][f] = _field_errors(m32.download(f), ref)
        out["fp64_floor"][f] = _field_errors(mf.download(f), ref)
    rounded = dict(arrays)
    for k in ("dem", "water_init", "mannings"):
        if rounded.get(k) is not None:
            rounded[k] = np.asarray(arrays[k], dtype=np.float32).astype(np.float64)
    mr = Model(None, None, precision=hc.HC_FP64, arrays=rounded, params=params, device_loop=True,
               use_graph=True)
    mr.iterate(nsteps)
    mr.synchronize()
    out["fp64_floor_f32_inputs"] = {f: _field_errors(mr.download(f), m64.download(f))
                                    for f in ("water_depth", "qx", "qy")}
    out["fp64_floor_f32_inputs"]["what"] = ("the FP64 form from the DEM, initial depth and "
                                            "Manning n rounded to float32, against the FP64 form")
    mr.close()
    vol32 = float(m32.download("water_depth")[1:-1, 1:-1].sum())
    vol64 = float(m64.download("water_depth")[1:-1, 1:-1].sum())
    out["volume_rel"] = abs(vol32 - vol64) / vol64
    h, hf = out["fields"]["water_depth"], out["fp64_floor"]["water_depth"]
    out["fp64_floor"]["what"] = "the FP64 form from the DEM plus 1e-8 m of noise, against the FP64 form"
    out["meets_max_1e-4"] = bool(h["max_rel"] <= 1e-4)
    out["within_fp64_floor"] = bool(all(h[k] <= hf[k] for k in ("max_rel", "l2_rel", "frac_cells_over_1e-4")))
    hr = out["fp64_floor_f32_inputs"]["water_depth"]
    out["within_fp64_floor_f32_inputs"] = bool(all(h[k] <= hr[k] for k in ("max_rel", "l2_rel", "frac_cells_over_1e-4")))
    mf.close()
    return out


def run_ensemble(steps: int, warmup: int, n: int = 64):
    """The batched ensemble: n independent catchments (side lengths
    log-uniform in [128, 1024] cells at 5 m, own DEM, n and rain), FP32 with
    fixed-point elevations, each with its own adaptive dt, stepped together from a
    captured graph; started fully wet (0.1 m) so that every cell is active."""
    from hail_caesar_amd import hc, inputs
    from hail_caesar_amd.model import Ensemble, Model
    specs = inputs.ensemble_members(n, hours=6, wet_depth=0.1)
    cells = int(sum(a["dem"].shape[0] * a["dem"].shape[1] for a, _ in specs))
    ms = [Model(None, None, precision=hc.HC_FP32, arrays=a, params=p, device_loop=True,
                elev_fixed=True) for a, p in specs]
    ens = Ensemble(ms)
    ens.iterate(warmup)
    rows, blocks, _ = ens.tune()
    ens.iterate(2)
    t = time.perf_counter()
    ens.run(steps)
    wall = time.perf_counter() - t
    ens.collect(steps)
    kus = ens.bench_fused(10)
    # the walking ("persistent") grids the tuner weighed, at the tuned tile rows:
    # blocks that each walk a contiguous run of the tile list
    walk = {}
    for per_cu in (2, 4, 8):
        ens.set_tiles(rows, 256 * per_cu)
        walk[str(256 * per_cu)] = round(ens.bench_fused(10), 2)
    ens.set_tiles(rows, blocks)
    sizes = sorted(a["dem"].shape[0] for a, _ in specs)
    out = {"members": n, "cells": cells, "side_min": sizes[0], "side_max": sizes[-1],
           "steps": steps, "value": round(cells * steps / wall, 1), "unit": "cell-steps/s",
           "ms_per_step": round(1000 * wall / steps, 5), "fused_kernel_us": round(kus, 2),
           "fused_gbs": round(bytes_per_cell_step("fp32", spatial_n=False) * cells / kus / 1e3, 1),
           "tile_rows": rows, "fused_grid_blocks": blocks,
           "walking_grid_us": walk,
           "sim_minutes_range": [round(min(m.scalar("cycle") for m in ms), 3),
                                 round(max(m.scalar("cycle") for m in ms), 3)]}
    ens.close()
    for m in ms:
        m.close()
    return out


def run_adjoint(size: int = 512, steps: int = 2000, dt: float = 0.5, every: int = 100):
    """The adjoint workload (inputs.adjoint_case): size^2 cells at 5 m,
    fractal DEM (5 m relief, 1 % slope to the south edge), log-normal n, a
    uniform 0.1 m start, 8 rain blocks (2 x 4) under a varying rain rate,
    `steps` steps at a fixed dt; the loss is the outlet-discharge misfit
    against a target made with n perturbed by 10 % smooth noise.  Times one
    forward (loss) and one reverse sweep (gradients w.r.t. n and the 8
    multipliers, including the segment recomputation from the checkpoints)."""
    from hail_caesar_amd import hc, inputs
    case = inputs.adjoint_case(size=size, steps=steps)
    pad = inputs.pad
    p = hc.default_params(hflow_threshold=0.001, edgeslope=0.01, spatially_var_mannings=1)
    ctx = hc.Context(size, size, case["DX"], p, hc.HC_FP64)
    ctx.upload("elev", pad(case["dem"], -9999.0))
    ctx.upload("water_depth", pad(case["water"], 0.0))
    ctx.upload("mannings", pad(case["n_target"], 0.0))
    adj = hc.Adjoint(ctx, steps, dt, case["block"], case["rain"], checkpoint_every=every)
    _, qt = adj.forward(np.ones(8))
    adj.set_target(qt)
    ctx.upload("mannings", pad(case["n"], 0.0))
    adj.forward(np.ones(8))  # warm
    t0 = time.perf_counter()
    loss, _ = adj.forward(np.ones(8))
    fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    gn, gm = adj.backward()
    bwd = time.perf_counter() - t0
    out = {"grid": f"{size}x{size} at 5 m", "steps": steps, "dt_s": dt, "checkpoint_every": every,
           "rain_blocks": 8, "dtype": "f64", "loss": loss,
           "forward_ms_per_step": round(1000 * fwd / steps, 5),
           "reverse_ms_per_step": round(1000 * bwd / steps, 5),
           "adjoint_over_forward": round(bwd / fwd, 3),
           "grad_n_l2": float(np.sqrt((gn ** 2).sum())), "grad_mult": [float(x) for x in gm]}
    adj.close()
    ctx.close()
    return out


def erosion_accuracy_run(size: int, hours: int):
    """The flow + erosion configuration at its stated shape: from dry, the
    6 h storm, `hours` simulated, host-stepped as the reference's loop, in the
    FP64 form, the FP32 storage form and the FP64 form from its inputs (DEM,
    grain data) rounded to float32; per-field errors at the end and the outlet
    water and sediment series' errors over the run."""
    from hail_caesar_amd import hc, inputs
    from hail_caesar_amd.model import Model
    arrays, params = inputs.erosion_arrays(size, wet_start=False, hours=hours)
    rounded = dict(arrays)
    for k in ("dem", "graindata"):
        rounded[k] = np.asarray(arrays[k], dtype=np.float32).astype(np.float64)
    res = {}
    names = ("elev", "water_depth", "qx", "qy", "Vsusptot", "grain")
    for prec, name, arr in ((hc.HC_FP64, "fp64", arrays), (hc.HC_FP32, "fp32", arrays),
                            (hc.HC_FP64, "fp64_floor", rounded)):
        m = Model(None, None, precision=prec, arrays=arr, params=params)
        t0 = time.perf_counter()
        n = 0
        while True:
            tt, cy, dt = m.iterate(5000)
            n += len(tt)
            print(f"[erosion accuracy] {name}: {n} bodies, {m.scalar('cycle') / 60:.2f} h",
                  file=sys.stderr, flush=True)
            if m.finished:
                break
        res[name] = {"fields": {f: m.download(f) for f in names}, "series": m.series(), "bodies": n,
                     "wall_s": round(time.perf_counter() - t0, 1)}
        m.close()
    b = res["fp64"]["fields"]
    dem = arrays["dem"]
    change = float(np.abs(b["elev"][1:-1, 1:-1] - dem).max())

    def errs(r):
        a = r["fields"]
        e = {"elev": {"max_abs": float(np.abs(a["elev"] - b["elev"])[1:-1, 1:-1].max()),
                      "max_elevation_change": change}}
        for f in names[1:]:
            e[f] = _field_errors(a[f] if a[f].ndim == 2 else a[f].sum(axis=2),
                                 b[f] if b[f].ndim == 2 else b[f].sum(axis=2))
        return e

    def series_err(x, y, col):
        k = min(len(x), len(y))
        d = x[:k, col] - y[:k, col]
        ref = np.abs(y[:k, col]).max()
        return {"rows": int(k), "max_rel": float(np.abs(d).max() / ref) if ref > 0 else 0.0,
                "l2_rel": float(np.sqrt((d ** 2).sum() / (y[:k, col] ** 2).sum())) if ref > 0 else 0.0,
                "total": float(x[:k, col].sum()), "total_fp64": float(y[:k, col].sum())}
    sb = res["fp64"]["series"]
    out = {"workload": f"{size}^2 at 10 m, 9 fractions (seeded Dirichlet), Wilcock-Crowe, 1 m "
                       f"active layer, 6 h triangular storm peaking at 40 mm/h, dry start, {hours} h",
           "bodies": {k: v["bodies"] for k, v in res.items()},
           "wall_s": {k: v["wall_s"] for k, v in res.items()},
           "fp32": {"fields": errs(res["fp32"]),
                    "water_out": series_err(res["fp32"]["series"], sb, 1),
                    "sediment_out": series_err(res["fp32"]["series"], sb, 4)},
           "fp64_floor_f32_inputs": {"fields": errs(res["fp64_floor"]),
                                     "water_out": series_err(res["fp64_floor"]["series"], sb, 1),
                                     "sediment_out": series_err(res["fp64_floor"]["series"], sb, 4)},
           "note": "grain: the cells' summed active-layer grain; series columns: the .dat file's "
                   "water (2nd column) and sediment (5th) discharge per output interval"}
    return out


def run_erosion(size: int = 1024, steps: int = 200, warmup: int = 50, cpu: bool = True):
    """Flow + erosion (host-stepped as the reference's loop): size^2 at 10 m,
    fractal relief 20 m on a 1 % slope with a carved channel, 9 grain
    fractions (d = 0.0005 .. 0.128 m) in seeded Dirichlet surface proportions
    per cell, Wilcock-Crowe, 1 m active layer, a 6 h triangular storm peaking
    at 40 mm/h, from a uniform 0.2 m start so that flow and entrainment are
    active everywhere.  Times whole loop bodies (flow_route, depth_update,
    scan, erode at the reference's erosion interval, the hillslope processes
    on their triggers, water_flux_out, output) in the FP64 form and in the FP32
    storage form, then reports the FP32 form's per-field error against the FP64
    form after the same loop bodies."""
    from hail_caesar_amd import hc, inputs
    from hail_caesar_amd.model import Model
    arrays, params = inputs.erosion_arrays(size, wet_start=True, hours=24)
    dem = arrays["dem"]
    cells = size * size
    out = {"workload": f"{size}^2 at 10 m, 9 fractions (seeded Dirichlet proportions), "
                       "Wilcock-Crowe, 1 m active layer, 40 mm/h storm, wet start (0.2 m)",
           "steps": steps, "unit": "cell-steps/s"}
    fields = {}
    for name, prec in (("fp64", hc.HC_FP64), ("fp32", hc.HC_FP32)):
        m = Model(None, None, precision=prec, arrays=arrays, params=params)
        m.iterate(warmup)
        m.synchronize()
        t = time.perf_counter()
        m.iterate(steps)
        m.synchronize()
        wall = time.perf_counter() - t
        out[name] = {"value": round(cells * steps / wall, 1), "ms_per_step": round(1000 * wall / steps, 4),
                     "sim_minutes": round(m.scalar("cycle"), 3)}
        fields[name] = {f: m.download(f) for f in ("elev", "water_depth", "qx", "qy", "Vsusptot", "grain")}
        m.close()
        print(f"[bench] erosion {name}: {out[name]['value']:.4e} cell-steps/s, "
              f"{out[name]['ms_per_step']} ms/step", file=sys.stderr, flush=True)
    # the FP64 form's own response to 1e-8 m of DEM noise, after the same bodies
    noisy = dict(arrays)
    noise = 1e-8 * np.random.Generator(np.random.PCG64(3)).standard_normal(dem.shape)
    noisy["dem"] = np.where(dem > inputs.NODATA, dem + noise, dem)
    m = Model(None, None, precision=hc.HC_FP64, arrays=noisy, params=params)
    m.iterate(warmup + steps)
    fields["floor"] = {f: m.download(f) for f in ("elev", "water_depth", "qx", "qy", "Vsusptot", "grain")}
    m.close()
    # and from the inputs rounded to float32 (DEM, initial depth, grain data):
    # erode_mult = trunc(ERODEFACTOR / tempbmax) (LCM.cpp:2277-2294) is
    # discontinuous, and a sub-ulp perturbation of this size changes its
    # schedule within ~100 bodies, as the FP32 form's rounding does
    rounded = dict(arrays)
    for k in ("dem", "water_init", "graindata"):
        rounded[k] = np.asarray(arrays[k], dtype=np.float32).astype(np.float64)
    m = Model(None, None, precision=hc.HC_FP64, arrays=rounded, params=params)
    m.iterate(warmup + steps)
    fields["floor32"] = {f: m.download(f) for f in ("elev", "water_depth", "qx", "qy", "Vsusptot", "grain")}
    m.close()
    out["value"] = out["fp64"]["value"]
    out["dtype"] = "f64"
    b = fields["fp64"]
    change = float(np.abs(b["elev"][1:-1, 1:-1] - dem).max())

    def errs(a):
        e = {"elev": {"max_abs": float(np.abs(a["elev"] - b["elev"])[1:-1, 1:-1].max()),
                      "max_elevation_change": change}}
        for f in ("water_depth", "qx", "qy", "Vsusptot", "grain"):
            e[f] = _field_errors(a[f] if a[f].ndim == 2 else a[f].sum(axis=2),
                                 b[f] if b[f].ndim == 2 else b[f].sum(axis=2))
        return e
    out["fp32_error"] = {"after_loop_bodies": warmup + steps, "fields": errs(fields["fp32"]),
                         "fp64_floor": errs(fields["floor"]),
                         "fp64_floor_f32_inputs": errs(fields["floor32"]),
                         "note": "grain: the cells' summed active-layer grain; fp64_floor: the FP64 "
                                 "form from the DEM plus 1e-8 m of noise, against the FP64 form; "
                                 "fp64_floor_f32_inputs: the FP64 form from the DEM, initial depth "
                                 "and grain data rounded to float32"}
    if cpu:
        out["cpu_baseline"] = cpu_erosion_baseline(arrays, params, warmup)
        print(f"[bench] erosion cpu baseline: {out['cpu_baseline']['value']:.4e} cell-steps/s",
              file=sys.stderr, flush=True)
    return out


def cpu_erosion_baseline(arrays, params, warmup: int, budget_s: float = 15.0):
    """The oracle (FP64 OpenMP restatement, every core this process may use)
    on the same flow + erosion workload: whole loop bodies after `warmup`
    bodies, a bounded sample of about budget_s.  Beside it, the port's speed
    ratio to the compiled reference on the erosion run where the reference can
    be built (oracle/time_cpu.py --erosion, profiles/cpu_baseline_erosion_256.json)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from hail_caesar_amd import inputs
    from pyoracle import OracleModel, pad, params_from_dict
    threads = host_cores()
    dem = arrays["dem"]
    p = params_from_dict(params, dem.shape[0], dem.shape[1], float(arrays["DX"]), nthreads=threads,
                         grain_d=inputs.EROSION_GRAIN_D)
    o = OracleModel(p)
    o.field("elev")[...] = pad(dem, -9999.0)
    o.field("water_depth")[1:-1, 1:-1] = arrays["water_init"]
    o.set_rain(np.asarray(arrays["rain"], dtype=np.float64))
    o.load_grain_arrays(arrays["graindata"])
    o.L.orc_initialise(o.h)
    warmup = min(warmup, 3)  # the first body runs creep and global landsliding
    t = time.perf_counter()
    o.iterate(warmup)
    warm = time.perf_counter() - t
    # bodies with and without an erode call differ widely in cost: step
    # until the budget is spent
    n = 0
    t = time.perf_counter()
    while True:
        o.iterate(1)
        n += 1
        el = time.perf_counter() - t
        if el >= budget_s or n >= 2000:
            break
    cells = dem.shape[0] * dem.shape[1]
    out = {"value": round(cells * n / el, 1), "unit": "cell-steps/s", "cores": threads, "kind": "port",
           "sample": f"oracle (FP64 OpenMP restatement, {threads} threads) on the same "
                     f"{dem.shape[0]}x{dem.shape[1]} flow + erosion state: {n} loop bodies after "
                     f"{warmup} warm-up bodies ({warm:.1f} s), {el:.1f} s"}
    ref = os.path.join(ROOT, "profiles", "cpu_baseline_erosion_256.json")
    if os.path.exists(ref):
        with open(ref) as f:
            r = json.load(f)
        out["port_over_reference"] = round(r["port_over_reference"], 3)
        out["reference_estimate"] = round(out["value"] / r["port_over_reference"], 1)
        out["reference_timing"] = (
            f"the compiled reference (-O3 -march=native -fopenmp) and this port on the same "
            f"{r['workload']} with {r['host']['threads']} threads: "
            f"{r['runs']['reference_o3']['cell_steps_per_s']:.3e} and "
            f"{r['runs']['port']['cell_steps_per_s']:.3e} cell-steps/s")
    return out


def accuracy_run(size: int, hours: int):
    """The large single grid as specified: dry start, `hours` of rain, run to
    the end in the FP32 form (fixed-point elevations) and the FP64 form; per-field
    errors at the end and the outlet-hydrograph error over the run."""
    from hail_caesar_amd import hc, inputs
    from hail_caesar_amd.model import Model
    arrays, params = inputs.grid4096_arrays(size=size, wet_depth=None, hours=hours)
    out = {"workload": f"{size}^2 at 2 m, dry start, {hours} h of 15-min Gamma rain", "fields": {}}
    res = {}
    pert = dict(arrays)
    pert["dem"] = np.asarray(arrays["dem"], dtype=np.float32).astype(np.float64)
    for prec, name, arr in ((hc.HC_FP64, "fp64", arrays), (hc.HC_FP32, "fp32", arrays),
                            (hc.HC_FP64, "fp64_floor", pert)):
        m = Model(None, None, precision=prec, arrays=arr, params=params, device_loop=True,
                  use_graph=True, elev_fixed=prec == hc.HC_FP32)
        t0 = time.perf_counter()
        n = 0
        while True:
            tt, cy, dt = m.iterate(20000)
            n += len(tt)
            print(f"[accuracy] {name}: {n} steps, {m.scalar('cycle') / 60:.2f} h", file=sys.stderr,
                  flush=True)
            if m.finished:
                break
        res[name] = {"model": m, "series": m.series(), "steps": n,
                     "wall_s": time.perf_counter() - t0}
    a, b, fl = res["fp32"], res["fp64"], res["fp64_floor"]
    out["fp64_floor"] = {"what": "the FP64 form from the DEM rounded to float32, against the FP64 form",
                         "fields": {}}
    for f in ("water_depth", "qx", "qy"):
        ref = b["model"].download(f)
        out["fields"][f] = _field_errors(a["model"].download(f), ref)
        out["fp64_floor"]["fields"][f] = _field_errors(fl["model"].download(f), ref)

    def hyd(x, y):
        k = min(len(x), len(y))
        return {"rows": int(k), "max_rel": float(np.abs(x[:k] - y[:k]).max() / np.abs(y[:k]).max()),
                "l2_rel": float(np.sqrt(((x[:k] - y[:k]) ** 2).sum() / (y[:k] ** 2).sum())),
                "peak": float(x[:k].max()), "peak_fp64": float(y[:k].max())}
    out["hydrograph"] = hyd(a["series"][:, 1], b["series"][:, 1])
    out["fp64_floor"]["hydrograph"] = hyd(fl["series"][:, 1], b["series"][:, 1])
    out["steps"] = {"fp32": a["steps"], "fp64": b["steps"]}
    out["wall_s"] = {"fp32": round(a["wall_s"], 1), "fp64": round(b["wall_s"], 1)}
    vol = [float(r["model"].download("water_depth")[1:-1, 1:-1].sum()) for r in (a, b)]
    out["volume_rel"] = abs(vol[0] - vol[1]) / max(vol[1], 1e-300)
    return out


def host_cores() -> int:
    """Cores this process may run on: the CPU affinity set, capped by the
    cgroup CPU quota where one is set (cpu.max)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if quota != "max":
            n = min(n, max(1, int(int(quota) // int(period))))
    except (OSError, ValueError):
        pass
    return max(1, n)


def cpu_baseline(arrays, params, budget_s: float = 20.0):
    """The oracle (FP64 CPU restatement, OpenMP on every core this process
    may use) on the same 4096^2 state: timed over whole loop bodies after the
    active mask is built.  Beside it, the port's speed ratio to the compiled
    reference (-O3 -march=native -fopenmp), both timed on identical inputs
    where the reference can be built (oracle/time_cpu.py,
    profiles/cpu_baseline_4096.json)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pyoracle import OracleModel, pad, params_from_dict
    threads = host_cores()
    dem = arrays["dem"]
    p = params_from_dict(params, dem.shape[0], dem.shape[1], float(arrays["DX"]), nthreads=threads)
    o = OracleModel(p)
    o.field("elev")[...] = pad(dem, -9999.0)
    o.field("mannings")[1:-1, 1:-1] = arrays["mannings"]
    o.field("water_depth")[1:-1, 1:-1] = arrays["water_init"]
    o.set_rain(np.asarray(arrays["rain"], dtype=np.float64))
    o.L.orc_initialise(o.h)
    o.iterate(5)  # the active mask is first built at counter 5
    t = time.perf_counter()
    o.iterate(1)
    one = time.perf_counter() - t
    n = max(1, min(1000, int(budget_s / max(one, 1e-3))))
    t = time.perf_counter()
    o.iterate(n)
    el = time.perf_counter() - t
    cells = dem.shape[0] * dem.shape[1]
    out = {"value": cells * n / el, "unit": "cell-steps/s", "cores": threads, "kind": "port",
           "host_cpus": os.cpu_count(),
           "sample": f"oracle (FP64 OpenMP restatement, {threads} threads) on the same fully wet "
                     f"{dem.shape[0]}x{dem.shape[1]} state: {n} loop bodies after 6 warm-up bodies, "
                     f"{el:.1f} s"}
    ref = os.path.join(ROOT, "profiles", "cpu_baseline_4096.json")
    if os.path.exists(ref):
        with open(ref) as f:
            r = json.load(f)
        ratio = r["port_over_reference"]
        out["port_over_reference"] = round(ratio, 3)
        out["reference_estimate"] = round(out["value"] / ratio, 1)
        out["reference_timing"] = (
            f"the compiled reference (-O3 -march=native -fopenmp) and this port timed on identical "
            f"inputs with {r['host']['threads']} threads ({r['host']['cpu']}): "
            f"{r['runs']['reference_o3']['cell_steps_per_s']:.3e} and "
            f"{r['runs']['port']['cell_steps_per_s']:.3e} cell-steps/s; reference_estimate is the "
            "port's value here divided by that ratio")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=None,
                    help="grid side (default 4096; 1024 for --erosion-hours, the flow + erosion "
                         "configuration's stated shape)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["fp32", "fp64"], default=None)
    ap.add_argument("--no-graph", action="store_true", help="launch every kernel from the host")
    ap.add_argument("--tiles", default=None, metavar="RxS",
                    help="fixed fused-kernel tiling (rows x strip) instead of tuning")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the 64-catchment ensemble")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint timing")
    ap.add_argument("--no-erosion", action="store_true", help="skip the flow + erosion timing")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first run a plain device copy of MIB MiB (counter calibration)")
    ap.add_argument("--section", choices=["ensemble", "erosion", "adjoint"], default=None,
                    help="run only this part (its JSON object alone; for profiler runs)")
    ap.add_argument("--erosion-hours", type=int, default=0, metavar="H",
                    help="instead of the benchmark, run the flow + erosion configuration from dry "
                         "for H hours (size from --size, 1024 for the stated shape) in FP64, the "
                         "FP32 storage form and FP64 from float32-rounded inputs, and report "
                         "their differences")
    ap.add_argument("--accuracy-hours", type=int, default=0, metavar="H",
                    help="instead of the benchmark, run the large grid from dry for H hours in "
                         "FP32 and FP64 and report their differences")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU row-strip decomposition is a design, not a build).", file=sys.stderr)
        sys.exit(2)
    load_package()
    if args.accuracy_hours:
        print(json.dumps({"accuracy": accuracy_run(args.size or 4096, args.accuracy_hours)}))
        return
    if args.erosion_hours:
        print(json.dumps({"erosion_accuracy": erosion_accuracy_run(args.size or 1024, args.erosion_hours)}))
        return
    if args.section == "ensemble":
        print(json.dumps({"ensemble": run_ensemble(args.steps, args.warmup)}))
        return
    if args.section == "erosion":
        print(json.dumps({"erosion": run_erosion(steps=args.steps, warmup=args.warmup,
                                                 cpu=not args.no_cpu)}))
        return
    if args.section == "adjoint":
        print(json.dumps({"adjoint": run_adjoint()}))
        return
    if args.copy_probe:
        copy_probe(args.copy_probe)
    from hail_caesar_amd import inputs
    with open(BASELINE) as f:
        base = json.load(f)
    if args.size is None:
        args.size = 4096
    arrays, params = inputs.grid4096_arrays(size=args.size)
    cells = args.size * args.size
    res = {}
    for prec in (["fp32", "fp64"] if args.only is None else [args.only]):
        r = run_gpu(prec, arrays, params, args.steps, args.warmup, use_graph=not args.no_graph,
                    tiles=tuple(int(v) for v in args.tiles.split("x")) if args.tiles else None)
        bpc = bytes_per_cell_step(prec)
        r["cell_steps_per_s"] = cells * args.steps / r["wall_s"]
        r["ms_per_step"] = 1000.0 * r["wall_s"] / args.steps
        r["kernel_gbs"] = bpc * cells / (r["kernel_ms"] * 1e-3) / 1e9
        r["bytes_per_cell_step"] = bpc
        res[prec] = r
        print(f"[bench] {prec}: {r['cell_steps_per_s']:.4e} cell-steps/s, {r['ms_per_step']:.4f} ms/step, "
              f"fused kernel {r['kernel_ms']*1e3:.1f} us = {r['kernel_gbs']:.0f} GB/s, "
              f"active {r['active_fraction']:.3f}, tiles {r['tile_rows']}x{r['tile_strip']}", file=sys.stderr, flush=True)
    head = res.get("fp32") or res.get("fp64")
    prec = "fp32" if "fp32" in res else "fp64"
    traffic = None
    if os.path.exists(TRAFFIC):
        with open(TRAFFIC) as f:
            t = json.load(f)
        # HBM bytes per launch from the FETCH_SIZE / WRITE_SIZE counter runs,
        # scaled by the same counters' reading of a plain copy of known size
        traffic = t.get(f"{prec}_R{head['tile_rows']}_S{head['tile_strip']}_bytes_per_launch",
                        t.get(f"{prec}_bytes_per_launch"))
    roofline = {"bound": "hbm", "achieved": round(head["kernel_gbs"], 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(head["kernel_gbs"] / HBM_PEAK_GBS, 4),
                "traffic": traffic,
                "kernel": "k_fused<3,R> (fused flow_route+depth_update, FP32 state, fixed-point z)"
                if prec == "fp32" else "k_fused<0,R> (fused flow_route+depth_update, FP64)",
                "bytes_per_cell_step": head["bytes_per_cell_step"], "cells_per_launch": cells,
                "kernel_us": round(head["kernel_ms"] * 1e3, 2),
                "tile_loop_us": head["loop_tile_us"], "tile_rows": head["tile_rows"],
                "tile_strip": head["tile_strip"]}
    out = {
        "metric": base["metric"], "value": round(head["cell_steps_per_s"], 1), "unit": "cell-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if prec == "fp32" else "f64", "data": "synthetic",
        "config": {"workload": f"{args.size}^2 grid at 2 m, fractal DEM (30 m) + D8-carved drainage network, "
                               "log-normal Manning n, 15-min Gamma rain, flow only, fully wet "
                               "(0.2 m initial depth); one step = one loop body, device-resident",
                   "cells": cells, "active_fraction": round(head["active_fraction"], 4)},
        "roofline": roofline,
    }
    if "fp32" in res and "fp64" in res:
        f32, f64 = res["fp32"], res["fp64"]
        out["fp64"] = {"value": round(f64["cell_steps_per_s"], 1), "ms_per_step": round(f64["ms_per_step"], 5),
                       "kernel_us": round(f64["kernel_ms"] * 1e3, 2),
                       "kernel_gbs": round(f64["kernel_gbs"], 1),
                       "frac": round(f64["kernel_gbs"] / HBM_PEAK_GBS, 4)}
        out["fp32_over_fp64"] = round(f32["cell_steps_per_s"] / f64["cell_steps_per_s"], 3)
        out["fp32_error"] = fp32_error(f32, f64, arrays, params)
    if not args.no_ensemble and args.only is None:
        out["ensemble"] = run_ensemble(args.steps, args.warmup)
        print(f"[bench] ensemble: {out['ensemble']['value']:.4e} cell-steps/s, "
              f"{out['ensemble']['ms_per_step']:.4f} ms/step, fused {out['ensemble']['fused_kernel_us']} us",
              file=sys.stderr, flush=True)
    if not args.no_erosion and args.only is None:
        out["erosion"] = run_erosion(cpu=not args.no_cpu)
    if not args.no_adjoint and args.only is None:
        out["adjoint"] = run_adjoint()
        print(f"[bench] adjoint: forward {out['adjoint']['forward_ms_per_step']} ms/step, reverse "
              f"{out['adjoint']['reverse_ms_per_step']} ms/step, ratio {out['adjoint']['adjoint_over_forward']}",
              file=sys.stderr, flush=True)
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(arrays, params)
        except Exception as e:  # the baseline is reported, never required
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Test infrastructure: the oracle under AddressSanitizer + UndefinedBehavior-
Sanitizer and under ThreadSanitizer (its OpenMP loops on several threads).

Builds oracle/_san/liboracle_{asan,tsan}.so (oracle/Makefile `san`), then
runs each workload below in a child Python with the sanitizer runtime
preloaded and HC_ORACLE_SO pointing at the build, and in a child with the
plain liboracle.so.  A sanitizer report ends the child with an error; the
final fields, the per-iteration series and the grain-row numbers of every
sanitized run must equal the plain run's bit for bit.

  flow64      the flow-only workload, 64^2, 1 h (about 1300 iterations)
  erosion48   the erosion run, 48^2, 1 h, half the cells in the grain data
              file (rows made as it goes; creep and landslides fire)
  reach40     reach mode with sediment (erosion), 40^2, 300 iterations
  rain3       spatially complex rainfall, 3 rain cells, 40^2, 1 h

Usage: python oracle/sanitize.py [--only asan|tsan] [--threads 4]
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SAN = os.path.join(HERE, "_san")
LLVM = os.environ.get("LLVM", "/opt/rocm/llvm")
WORKLOADS = ("flow64", "erosion48", "reach40", "rain3")


def _package():
    import importlib.util
    pkg = os.path.join(ROOT, "hail-caesar_amd")
    spec = importlib.util.spec_from_file_location(
        "hail_caesar_amd", os.path.join(pkg, "__init__.py"), submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["hail_caesar_amd"] = mod
    spec.loader.exec_module(mod)
    import hail_caesar_amd.inputs as inputs
    return inputs


def _digest(a) -> str:
    import numpy as np
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:16]


def child(name: str, threads: int) -> dict:
    """One workload on the oracle the environment selects; digests."""
    sys.path.insert(0, HERE)
    import numpy as np
    inputs = _package()
    from pyoracle import OracleModel
    with tempfile.TemporaryDirectory() as tmp:
        steps = 10 ** 7
        if name == "flow64":
            w = inputs.flow_only_256(os.path.join(tmp, "w"), size=64, hours=1)
        elif name == "erosion48":
            w = inputs.erosion_run(os.path.join(tmp, "w"), grain_cover=0.5)
        elif name == "reach40":
            w = inputs.reach_inputs(os.path.join(tmp, "w"), erosion=True, sed_scale=50.0)
            steps = 300
        else:
            w = inputs.complex_rain(os.path.join(tmp, "w"), 3)
        o = OracleModel.from_workload(w, nthreads=threads)
        tt, cy, dt = o.iterate(steps)
        out = {"steps": len(tt), "series": _digest(np.stack([tt, cy, dt]))}
        fields = ["water_depth", "qx", "qy", "elev"]
        if w.params.get("hydro_model_only", "yes") == "no":
            fields += ["Vsusptot", "grain", "strata"]
            out["index"] = _digest(o.field_int("index"))
        for f in fields:
            out[f] = _digest(o.field(f))
    # which oracle build and sanitizer runtime the process really mapped
    maps = open("/proc/self/maps").read()
    out["mapped"] = sorted({os.path.basename(ln.split()[-1]) for ln in maps.splitlines()
                            if ln.split()[-1].startswith("/") and
                            ("liboracle" in ln or "asan" in ln or "tsan" in ln or "ubsan" in ln
                             or "archer" in ln)})
    return out


def sanitizer_env(kind: str, threads: int) -> dict:
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    if kind == "asan":
        rt = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True,
                            text=True).stdout.strip()
        env["LD_PRELOAD"] = rt
    elif kind == "tsan":
        rt = subprocess.run([os.path.join(LLVM, "bin", "clang++"),
                             "-print-file-name=libclang_rt.tsan-x86_64.so"], capture_output=True,
                            text=True).stdout.strip()
        env["LD_PRELOAD"] = rt
        env["OMP_TOOL_LIBRARIES"] = os.path.join(LLVM, "lib", "libarcher.so")
    if kind != "plain":
        env["HC_ORACLE_SO"] = os.path.join(SAN, f"liboracle_{kind}.so")
        # CPython's own allocations are not the oracle's: leaks are not checked
        env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=0:halt_on_error=1"
        env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
        env["TSAN_OPTIONS"] = "halt_on_error=1:report_signal_unsafe=0:ignore_noninstrumented_modules=1"
    return env


def run_child(name: str, kind: str, threads: int) -> dict:
    env = sanitizer_env(kind, threads)
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--child", name,
                        "--threads", str(threads)], env=env, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{kind} {name}: exit {r.returncode}\n{r.stderr[-4000:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])


CANARY = r"""
#include <cstdlib>
extern "C" double canary_race(int n) {  // unsynchronised += from every thread
  double s = 0;
#pragma omp parallel for
  for (int i = 0; i < n; ++i) s += i;
  return s;
}
extern "C" int canary_overflow(int n) {  // reads one past a heap block
  int* a = static_cast<int*>(std::malloc(sizeof(int) * n));
  for (int i = 0; i < n; ++i) a[i] = i;
  volatile int v = a[n];
  std::free(a);
  return v;
}
"""


def canary(kind: str, threads: int) -> bool:
    """Whether the sanitizer set-up reports a planted mistake: a data race
    (tsan) or a heap overrun (asan) in a two-function library built the same
    way as the oracle's sanitizer build.  Guards against a set-up that runs
    but checks nothing."""
    os.makedirs(SAN, exist_ok=True)
    src = os.path.join(SAN, "canary.cpp")
    with open(src, "w") as f:
        f.write(CANARY)
    so = os.path.join(SAN, f"libcanary_{kind}.so")
    if kind == "asan":
        cmd = ["g++", "-O1", "-g", "-fPIC", "-shared", "-fopenmp", "-fsanitize=address,undefined", src, "-o", so]
    else:
        cmd = [os.path.join(LLVM, "bin", "clang++"), "-O1", "-g", "-fPIC", "-shared", "-fopenmp",
               "-fsanitize=thread", f"-Wl,-rpath,{LLVM}/lib", src, "-o", so]
    subprocess.run(cmd, check=True)
    env = sanitizer_env(kind, threads)
    call = "canary_race(100000)" if kind == "tsan" else "canary_overflow(16)"
    r = subprocess.run([sys.executable, "-c", f"import ctypes; ctypes.CDLL({so!r}).{call}"],
                       env=env, capture_output=True, text=True)
    word = "data race" if kind == "tsan" else "heap-buffer-overflow"
    return r.returncode != 0 and word in r.stderr


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--child", default=None)
    ap.add_argument("--threads", type=int, default=4)
    ap.add_argument("--only", default=None, choices=["asan", "tsan"])
    args = ap.parse_args()
    if args.child:
        print(json.dumps(child(args.child, args.threads)), flush=True)
        return
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so", "san"], check=True)
    kinds = [args.only] if args.only else ["asan", "tsan"]
    bad = 0
    for kind in kinds:
        ok = canary(kind, args.threads)
        bad += not ok
        print(f"canary     {kind}: {'planted mistake reported' if ok else 'PLANTED MISTAKE NOT REPORTED'}",
              flush=True)
    for name in WORKLOADS:
        plain = run_child(name, "plain", args.threads)
        for kind in kinds:
            got = run_child(name, kind, args.threads)
            mapped = got.pop("mapped")
            assert f"liboracle_{kind}.so" in mapped and any(kind in m for m in mapped if "oracle" not in m), mapped
            plain.pop("mapped", None)
            same = got == plain
            bad += not same
            print(f"{name:10s} {kind}: {'clean, results identical' if same else 'RESULTS DIFFER'}"
                  f" ({plain['steps']} iterations)", flush=True)
    sys.exit(1 if bad else 0)


if __name__ == "__main__":
    main()

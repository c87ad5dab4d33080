"""Build libhc.so (HIP kernels + C-ABI) for gfx950 with hipcc, in-tree.

Usage: python hail-caesar_amd/build.py [--jobs N]
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libhc.so")
SOURCES = ["hc_capi.hip", "hc_hydro.hip", "hc_erode.hip", "hc_loop.hip", "hc_adjoint.hip", "hc_state.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Strict IEEE FP64 (no contraction); packed FP32 instructions disabled.
FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-I", os.path.join(HERE, "..", "include")]


def compile_one(src: str) -> str:
    os.makedirs(OBJ, exist_ok=True)
    obj = os.path.join(OBJ, src.replace(".hip", ".o"))
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    if r.stderr.strip():
        sys.stderr.write(r.stderr)
    return obj


HOST_SOURCES = ["driver/hc_model.cpp"]
CXX = os.environ.get("CXX", "g++")
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra"]
BIN = os.path.join(HERE, "bin", "hail-caesar-mi355x")
# The same driver with its host code (run loop, parsers, writers,
# checkpoints) under AddressSanitizer and UndefinedBehaviorSanitizer, for
# tests; the device code in libhc.so is the ordinary build.
BIN_ASAN = os.path.join(HERE, "bin", "hail-caesar-mi355x-asan")
ASAN_FLAGS = ["-O1", "-g", "-fno-omit-frame-pointer", "-std=c++17", "-ffp-contract=off",
              "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"]


def compile_host(src: str) -> str:
    os.makedirs(OBJ, exist_ok=True)
    obj = os.path.join(OBJ, os.path.basename(src).replace(".cpp", ".o"))
    cmd = [CXX, *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int = 4) -> str:
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(compile_one, s) for s in SOURCES]
        futs += [ex.submit(compile_host, s) for s in HOST_SOURCES]
        objs = [f.result() for f in futs]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # command-line driver (src/main.cpp counterpart), linked against libhc.so
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    cmd = [CXX, *HOST_FLAGS, os.path.join(CSRC, "driver", "hc_main.cpp"), "-o", BIN,
           "-L" + HERE, "-lhc", "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver link failed:\n{r.stdout}\n{r.stderr}")
    cmd = [CXX, *ASAN_FLAGS, os.path.join(CSRC, "driver", "hc_main.cpp"),
           os.path.join(CSRC, "driver", "hc_model.cpp"), "-o", BIN_ASAN, "-L" + HERE, "-lhc",
           "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"sanitizer driver link failed:\n{r.stdout}\n{r.stderr}")
    return LIB


if __name__ == "__main__":
    jobs = int(sys.argv[sys.argv.index("--jobs") + 1]) if "--jobs" in sys.argv else 4
    print(build(jobs))

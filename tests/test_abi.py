"""The C-ABI boundary on the CPU: the shared library loads without a GPU and
exports every function the public headers declare, and the Python bindings
cover them all.  No compute call is made here."""
import ctypes
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "hail-caesar_amd")
LIB = os.path.join(PKG, "libhc.so")
HEADERS = [os.path.join(ROOT, "include", h) for h in ("hc.h", "hc_model.h")]
DECL = re.compile(r"^[A-Za-z_][\w ]*[\s*](hcm?_\w+)\(", re.M)


def declared():
    names = []
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)  # drop comments
        text = re.sub(r"//[^\n]*", "", text)
        names += DECL.findall(text)
    return sorted(set(names))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        subprocess.run([sys.executable, os.path.join(PKG, "build.py")], check=True)
    return ctypes.CDLL(LIB)


def test_headers_declare_the_boundary():
    names = declared()
    for must in ("hc_create", "hc_destroy", "hc_upload", "hc_download", "hc_flow_route",
                 "hc_depth_update", "hc_scan_area", "hc_water_flux_out", "hc_add_point_inputs",
                 "hc_erode", "hc_loop_begin", "hc_loop_run", "hc_loop_sync", "hcm_create",
                 "hcm_iterate"):
        assert must in names, must


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, f"declared in include/ but not exported: {missing}"


def test_bindings_cover_the_headers(lib):
    from hail_caesar_amd import hc
    bound = set(hc.SYMBOLS) | set(hc.MODEL_SYMBOLS)
    missing = [n for n in declared() if n not in bound]
    assert not missing, f"no ctypes binding for: {missing}"


def test_create_rejects_bad_arguments_without_a_device(lib):
    """Argument checks run before any device call: a null grid is refused
    with an error code, not a crash."""
    lib.hc_create.restype = ctypes.c_int
    lib.hc_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p, ctypes.c_void_p]
    h = ctypes.c_void_p()
    assert lib.hc_create(ctypes.byref(h), None, None) < 0

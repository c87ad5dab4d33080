import importlib.util
import os
import sys

try:
    # torch first: it brings its own HIP runtime libraries, and loading them
    # after libhc.so has initialised the system HIP runtime aborts the process
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def load_package():
    if "hail_caesar_amd" in sys.modules:
        return sys.modules["hail_caesar_amd"]
    pkg = os.path.join(ROOT, "hail-caesar_amd")
    spec = importlib.util.spec_from_file_location(
        "hail_caesar_amd", os.path.join(pkg, "__init__.py"), submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["hail_caesar_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


load_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libhc.so)")

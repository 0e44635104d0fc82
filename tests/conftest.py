import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "qd-openmp-cuda_amd")


def _load_package():
    if "qd_openmp_cuda_amd" in sys.modules:
        return sys.modules["qd_openmp_cuda_amd"]
    spec = importlib.util.spec_from_file_location(
        "qd_openmp_cuda_amd", os.path.join(PKG_DIR, "__init__.py"),
        submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["qd_openmp_cuda_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


_load_package()
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")

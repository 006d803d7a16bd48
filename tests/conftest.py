"""Shared test set-up: load the product package (its directory name has a
hyphen, so it is loaded by path as ``pstd_amd``) and the oracle binding."""
import importlib.util
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
PKG_DIR = ROOT / "pstd-scalarwavequantumscattering_amd"


def load_pkg():
    if "pstd_amd" in sys.modules:
        return sys.modules["pstd_amd"]
    spec = importlib.util.spec_from_file_location(
        "pstd_amd", PKG_DIR / "__init__.py", submodule_search_locations=[str(PKG_DIR)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["pstd_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
load_pkg()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")
    config.addinivalue_line("markers", "slow: longer CPU test")

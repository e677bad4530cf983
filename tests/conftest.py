"""Test configuration: the `gpu` marker, import paths, and in-tree builds of the oracle and library
when they are missing (a fresh checkout keeps built files out of history)."""
import importlib
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "hahenn-_amd")
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")
    oracle_so = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
    if not os.path.exists(oracle_so):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, capture_output=True)
    lib_so = os.path.join(PKG_DIR, "libsuiri_amd.so")
    if not os.path.exists(lib_so) and os.path.exists("/opt/rocm/bin/hipcc"):
        subprocess.run(["make", "-C", PKG_DIR, "-j8", "libsuiri_amd.so"], check=True, capture_output=True)


@pytest.fixture(scope="session")
def suiri():
    return importlib.import_module("hahenn-_amd")


@pytest.fixture(scope="session")
def scene_mod():
    return importlib.import_module("hahenn-_amd.scene")

"""Shared test set-up: module loading, markers, GPU context.

`-m "not gpu"` runs the oracle pinning tests, host-logic tests and the ABI
checks; `-m gpu` runs the parity tests, which call libmodlin.so through its C
ABI on a gfx950 device and compare with the oracle.
"""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "modlin-rs_amd")
sys.path.insert(0, os.path.join(ROOT, "tests"))


def _load(name, path):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


modlin = _load("modlin", os.path.join(PKG, "modlin.py"))
workloads = _load("modlin_workloads", os.path.join(PKG, "workloads.py"))
oracle = _load("modlin_oracle", os.path.join(ROOT, "oracle", "oracle.py"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (parity tests through the C ABI)")


@pytest.fixture(scope="session")
def ctx():
    c = modlin.Context(0)
    yield c
    c.close()

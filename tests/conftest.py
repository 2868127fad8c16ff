"""Shared fixtures: paths, the oracle, the engine bindings and GPU detection.

Tests marked ``gpu`` need a visible HIP device and the built engine library;
everything else runs on the CPU (the oracle, the host-side table builder of
the engine, the LXCat reader and the ABI surface).
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "loki-mc_amd", "python"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import ekmc  # noqa: E402
import oracle_py  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built libekmc.so")
    config.addinivalue_line("markers", "slow: longer statistical runs")


def _has_gpu():
    try:
        n = ekmc.C.c_int(0)
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        if hip.hipGetDeviceCount(ctypes.byref(n)) != 0:
            return False
        return n.value > 0
    except OSError:
        return False


HAS_GPU = _has_gpu()


def pytest_collection_modifyitems(config, items):
    if HAS_GPU:
        return
    skip = pytest.mark.skip(reason="no HIP device visible")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def root():
    return ROOT


@pytest.fixture(scope="session")
def cfg1():
    return ekmc.workload("cfg1_atomic40_dc")


def random_electrons(rng, n, emin=0.01, emax=50.0):
    e = rng.uniform(emin, emax, n)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    v = d * np.sqrt(2 * e * ekmc.CHARGE / ekmc.EMASS)[:, None]
    return v, 0.5 * ekmc.EMASS * (v * v).sum(1) / ekmc.CHARGE

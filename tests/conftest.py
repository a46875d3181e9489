"""Shared fixtures: the product package (ctypes over the HIP/host libraries)
and the CPU oracle.  GPU tests carry the `gpu` marker."""
import importlib
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PKG_NAME = "real-time-planetary-impacts-mps_amd"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def sph():
    return importlib.import_module(PKG_NAME)


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle


@pytest.fixture(scope="session")
def tables256(sph):
    return sph.make_eos_tables(sph.default_params(), 256)


_IC_CACHE = {}


@pytest.fixture(scope="session")
def impact(sph, tables256):
    """impact(n, seed, headon=False) -> initial conditions (cached)."""
    def make(n, seed, headon=False):
        key = (n, seed, headon)
        if key not in _IC_CACHE:
            prm = sph.default_params()
            _IC_CACHE[key] = sph.make_impact(prm, sph.impact_config(n, seed, headon=headon), tables256)
        return {k: v.copy() for k, v in _IC_CACHE[key].items()}
    return make

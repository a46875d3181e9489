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


@pytest.fixture(scope="session")
def demo_ic_raw():
    """The reference's demo initial conditions (SPH/demo_impact_n40.hdf5,
    10,766 particles), converted by tests/golden/make_demo_ic.py."""
    d = np.load(os.path.join(ROOT, "tests", "golden", "demo_impact_n40.npz"))
    return {k: d[k] for k in d.files}


@pytest.fixture
def demo_ic(demo_ic_raw):
    """Demo IC in the sph_ic_view layout (16-byte float3 stride)."""
    n = demo_ic_raw["mass"].shape[0]
    ic = {k: v.copy() for k, v in demo_ic_raw.items() if k not in ("pos", "vel")}
    for k in ("pos", "vel"):
        a = np.zeros((n, 4), np.float32)
        a[:, :3] = demo_ic_raw[k]
        ic[k] = a
    return ic

"""Shared fixtures.  GPU tests are marked ``gpu``; everything else runs on CPU."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "regularized-stokesian-dynamics_amd")
ORACLE = os.path.join(ROOT, "oracle")
for p in (PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)

DATA = os.path.join(PKG, "data")
LUBDAT = os.path.join(DATA, "lubdat")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built librsd_amd.so")


@pytest.fixture(scope="session")
def oracle():
    from rsd_oracle import Oracle
    o = Oracle()
    o.load_tables(LUBDAT)
    return o


@pytest.fixture(scope="session")
def reference():
    from rsd_oracle import Reference
    try:
        return Reference(lubdat_parent=DATA)
    except FileNotFoundError as e:
        pytest.skip(f"flang-built reference unavailable: {e}")


@pytest.fixture(scope="session")
def rsd():
    import rsd_amd
    rsd_amd.init(0)
    return rsd_amd


def relerr(a, b):
    """max |a-b| / max |a| (the parity metric used throughout)."""
    a = np.asarray(a)
    b = np.asarray(b)
    scale = np.max(np.abs(a))
    return float(np.max(np.abs(a - b)) / (scale if scale > 0 else 1.0))

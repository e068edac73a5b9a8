import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libicflow_hip.so on the device)")


def rel_err(got, ref):
    """Normwise relative error max|got - ref| / max|ref| (entries of an assembled array can
    cancel to ~0, so each array is measured against its own largest entry)."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    scale = np.max(np.abs(ref)) if ref.size else 0.0
    if scale == 0.0:
        return float(np.max(np.abs(got))) if got.size else 0.0
    return float(np.max(np.abs(got - ref)) / scale)


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle as o
    o.load()
    return o


@pytest.fixture(scope="session")
def lib():
    from icflow_amd import _lib
    return _lib.load()


@pytest.fixture(scope="session")
def specimen_consts(lib):
    from icflow_amd import _lib
    return _lib.consts_from_prms(**_lib.SPECIMEN_PRMS)

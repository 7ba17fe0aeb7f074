"""Test configuration: the `gpu` marker (tests that need an MI355X), shared
helpers, and a guard that the native libraries are built."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs a HIP GPU (MI355X); run with -m gpu")


def have_gpu():
    try:
        import mbsolve_amd.lib as mb
        return mb.device_count() > 0
    except Exception:
        return False


def rel_err(a, b):
    """max |a - b| / max |b| (relative error of a record against the
    reference, as defined for this project's measurements)."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    scale = np.max(np.abs(b))
    if scale == 0.0:
        return float(np.max(np.abs(a - b)))
    return float(np.max(np.abs(a - b)) / scale)


@pytest.fixture(scope="session")
def gpu():
    if not have_gpu():
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    return True

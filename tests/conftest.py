import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
FIXTURE_GSD = os.path.join(GOLDEN, "polydisperse12.gsd")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


def scaled_error(got, ref, absum, floor=1e-6):
    """Per-component error scaled by max(|T_ref|, floor * S), S = sum of |contributions|.

    The floor keeps components whose contributions cancel to ~0 from being
    judged against a relative error of their (rounding-level) value."""
    scale = np.maximum(np.abs(ref), floor * absum)
    scale = np.where(scale > 0, scale, 1.0)
    return np.abs(got - ref) / scale


@pytest.fixture(scope="session")
def gpu():
    from pyglasstools_amd import engine

    n = engine.device_count()
    if n < 1:
        pytest.fail("GPU test run without a visible HIP device")
    return n

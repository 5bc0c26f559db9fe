import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "mumax-plus_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity tests of the HIP path")
    config.addinivalue_line("markers", "slow: long-running oracle checks")


@pytest.fixture
def rng():
    return np.random.default_rng(1234)


def rel_max_err(result, wanted):
    """max |result - wanted| / max |wanted| over all cells and components."""
    result = np.asarray(result, dtype=np.float64)
    wanted = np.asarray(wanted, dtype=np.float64)
    scale = np.abs(wanted).max()
    return float(np.abs(result - wanted).max() / (scale if scale > 0 else 1.0))

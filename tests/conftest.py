import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(TESTS, "golden")

EPS2 = float(np.float32(0.1) * np.float32(0.1))  # softeningSquared as the reference computes it


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: full-size (N = 98304) checks")


@pytest.fixture(scope="session")
def oracle():
    from oracle_lib import oracle as _o
    return _o()


@pytest.fixture(scope="session")
def ref_cpu():
    from oracle_lib import ref_cpu as _r
    return _r()

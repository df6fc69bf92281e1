import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C-ABI on the device)")


@pytest.fixture(scope="session")
def tables():
    return dict(np.load(os.path.join(GOLDEN, "reference_tables.npz")))


@pytest.fixture(scope="session")
def goldens():
    return dict(np.load(os.path.join(GOLDEN, "scipy_goldens.npz")))


@pytest.fixture(scope="session")
def ke0(tables):
    return tables["ke0"]

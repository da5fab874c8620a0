import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C-ABI)")


def golden(*parts):
    return os.path.join(GOLDEN, *parts)


def read_ints(path):
    with open(path) as f:
        return np.array([int(t) for t in f.read().split()], dtype=np.int32)


@pytest.fixture(scope="session")
def ctx():
    from hethyppar_amd import hbs
    if hbs.device_count() < 1:
        pytest.fail("a GPU test ran without a visible HIP device")
    c = hbs.Context(0)
    yield c
    c.close()

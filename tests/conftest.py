import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, os.path.join(ROOT, "gaugefix-coulomb_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libgfc on the GPU)")
    config.addinivalue_line("markers", "slow: takes more than a few seconds")


def golden(name):
    path = os.path.join(GOLDEN, name)
    if not os.path.exists(path):
        pytest.skip(f"fixture {name} not generated")
    return np.load(path)


@pytest.fixture(scope="session")
def ctx():
    import gfc
    c = gfc.Context(0)
    yield c
    c.close()

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden", "data")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


def load_fixture(name):
    """Reference fixture extracted from test/data/*.jld (see tests/golden/make_fixtures.sh)."""
    return np.load(os.path.join(GOLDEN, name + ".npy"), allow_pickle=False)


@pytest.fixture(scope="session")
def fixture():
    return load_fixture

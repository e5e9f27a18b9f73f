import os
import sys

import numpy as np
import pytest

REPO = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if REPO not in sys.path:
    sys.path.insert(0, REPO)

GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def load_golden(name):
    path = os.path.join(GOLDEN, f"{name}.npz")
    if not os.path.exists(path):
        pytest.skip(f"golden fixture {name} missing")
    return np.load(path)


def lambert_dims(name, fx):
    """(npx, npy) that the reference passes to Calckvectors for the fixture's workload."""
    from ctemsoft_amd import ebsd, workloads
    base = name.split("_")[0] + "_" + name.split("_")[1]  # drop an energy tag (fe_c3w_10.0kV)
    return ebsd.lambert_dims(workloads.get(base).npx, int(fx["usehex"]))


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    oracle.lib()
    return oracle

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; runs the HIP kernels")
    config.addinivalue_line("markers", "slow: full-size workload (5M-point map)")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle as O

    O.lib()
    return O


@pytest.fixture(scope="session")
def full_wl():
    """The full-size workload: 5M-point map, ~29k-point scan."""
    from cuda_ndt_matcher_amd import synth

    return synth.make_workload("full", 0)


@pytest.fixture(scope="session")
def full_oracle(full_wl, oracle_mod):
    return oracle_mod.Grid(full_wl.map_points, 2.0)

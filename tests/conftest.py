import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libtiny_mc_gpu.so")
    config.addinivalue_line("markers", "slow: CPU test that takes more than a few seconds")


@pytest.fixture(scope="session")
def orc():
    import oracle_lib

    return oracle_lib.load()


@pytest.fixture(scope="session")
def tmc():
    import tiny_mc_amd

    tiny_mc_amd.load()
    return tiny_mc_amd


@pytest.fixture(scope="session")
def hw_tables(tmc):
    """The device's step length, sin/cos and sin(theta) for every uniform the
    walk can draw in the hardware-math mode (indexed by r >> 8), for the
    mirror."""
    return tmc.step_math_tables(tmc.MATH_HARDWARE)

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: large sizes")


@pytest.fixture(scope="session")
def cg():
    import crystalgraphics_amd

    return crystalgraphics_amd


@pytest.fixture(scope="session")
def oracle():
    from oracle import binding

    return binding

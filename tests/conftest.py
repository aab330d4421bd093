import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def ffi():
    from oracle import ffi as f
    f.lib()
    return f


@pytest.fixture(scope="session")
def ox():
    import ogs5_amd
    ogs5_amd.lib()
    return ogs5_amd

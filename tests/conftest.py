import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")


@pytest.fixture(scope="session")
def hiplib():
    """libcathode_hip.so, initialised on device 0 (GPU tests only)."""
    import importlib
    abi = importlib.import_module("02_cathode-08_modulerev_amd.abi")
    lib = abi.load()
    rc = lib.cathode_init(0)
    assert rc == 0, lib.cathode_last_error()
    yield lib
    lib.cathode_finalize()

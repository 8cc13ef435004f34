"""Test configuration: the `gpu` marker and the repository on sys.path."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and libmsim_hip.so")


@pytest.fixture(scope="session")
def oracle():
    from oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def reference():
    from oracle import Reference
    if not Reference.available():
        pytest.skip("oracle/_ref was not built (needs the reference tree)")
    return Reference()


@pytest.fixture(scope="session")
def hip():
    """The HIP backend on device 0. A GPU test must run the native library:
    a missing library or device is a failure, never a skip."""
    import microsimulator_amd as msa
    backend = msa.HipBackend(0)
    yield backend
    backend.close()

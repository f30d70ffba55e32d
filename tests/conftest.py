"""Shared fixtures. GPU tests are marked `gpu`; everything else runs on the CPU."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def aglib():
    import accelgraph_amd as ag
    return ag.lib()


@pytest.fixture(scope="session")
def device():
    import accelgraph_amd as ag
    if ag.Device.count() < 1:
        pytest.fail("GPU test run without a visible GPU")
    dev = ag.Device(0)
    yield dev
    dev.close()


@pytest.fixture(scope="session")
def rmat16():
    import accelgraph_amd as ag
    return ag.Graph.rmat(16)


@pytest.fixture(scope="session")
def rmat12():
    import accelgraph_amd as ag
    return ag.Graph.rmat(12)

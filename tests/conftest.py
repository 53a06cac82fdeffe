"""Shared fixtures.  GPU tests carry @pytest.mark.gpu; everything else runs
on the CPU (oracle, formats, generator, host logic, C ABI exports)."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def sw():
    import subway_amd
    return subway_amd


@pytest.fixture(scope="session")
def oracle():
    from tests import _oracle
    return _oracle


@pytest.fixture(scope="session")
def gpu(sw):
    if sw.device_count() < 1:
        pytest.fail("GPU test run without a visible HIP device")
    return True

"""pytest configuration: the `gpu` marker, import paths, device fixture."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def dev():
    """Open device 0 through libspx; fails (does not skip) when no GPU is present."""
    from streamparanumal_amd import device, ogs
    n = device.device_count()
    assert n > 0, "no HIP device visible to libspx"
    ogs.platform_open(0)  # opens the device through platform_t
    print(f"[gpu] device: {device.props()['name']} {device.props()['arch']}", flush=True)
    return device


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    pyoracle.lib()
    return pyoracle

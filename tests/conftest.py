import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device; calls the HIP path through the C ABI")


@pytest.fixture(scope="session")
def ctx():
    from helios_amd import native
    c = native.Context(0)
    yield c
    c.close()

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (runs the HIP library)")


@pytest.fixture(scope="session")
def handle():
    from steadyscan_amd import Handle
    h = Handle(0)
    yield h
    h.close()

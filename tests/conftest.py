import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libgst.so on cuda:0)")


@pytest.fixture(scope="session")
def solver():
    import gpu4gst_code_amd as gst
    s = gst.Solver(0)
    yield s
    s.close()

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("oracle", "mandelbrot-explorer_amd", "tests"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libmxhip.so")


@pytest.fixture(scope="session")
def ctx():
    import mxhip

    c = mxhip.Context(0)
    yield c
    c.close()

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (runs libqmb kernels)")


@pytest.fixture(scope="session")
def ctx():
    import qmbcuda_amd as q

    if q.device_count() < 1:
        pytest.fail("GPU test run without a visible GPU")
    c = q.Context(0)
    yield c
    c.close()

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")


@pytest.fixture(scope="session")
def ctx():
    """One HIP context (device 0, one stream) for every GPU test of the session."""
    from appfis_amd import hip

    c = hip.Context(0)
    print(f"[appfis] device: {c.name()}", flush=True)
    yield c
    c.close()

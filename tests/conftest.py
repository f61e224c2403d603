import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")
    config.addinivalue_line("markers", "slow: full-size workload")


@pytest.fixture(scope="session")
def ctx():
    """One library context for the whole GPU session (kernels are loaded once)."""
    import tensor_bf_main_amd as bfj

    c = bfj.Context(0)
    yield c
    c.close()

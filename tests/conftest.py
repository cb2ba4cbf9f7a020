import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libopsmi.so on the GPU)")


@pytest.fixture(scope="session")
def gpu_ops():
    """The backend initialised once per session on the GPU."""
    from ops_amd import ops

    ops.ops_init(diags=1)
    return ops

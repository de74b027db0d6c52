import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs the HIP path on an MI355X (gfx950)")


@pytest.fixture(scope="session")
def gpu():
    """The HIP device; a GPU test fails (never skips) when it is missing."""
    import torch
    assert torch.cuda.is_available(), "GPU test run without a visible HIP device"
    from superscanner_amd import _lib
    _lib.load()  # raises if libtsdf_hip.so was not built
    return torch.device("cuda", 0)

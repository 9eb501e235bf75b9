import os
import sys

import pytest

try:  # import early: the first import on a fresh machine is slow
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libtdvh.so)")


@pytest.fixture(scope="session")
def tv():
    import thundervision_amd.tdvh as tv
    tv.lib()  # fail loudly if the HIP library is missing
    return tv


@pytest.fixture(scope="session")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    return torch.device("cuda:0")

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "ewpt-gpu_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libewpt on the device)")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import cref
    cref.build()
    return cref


def pytest_sessionstart(session):
    """torch ships its own HIP runtime beside the system one that libewpt
    links.  Tests that park a 512^3 state in a torch device buffer need
    torch's runtime to come up before the library's; initialise it first
    (a no-op on a machine without a GPU)."""
    try:
        import torch
    except ImportError:
        return
    torch.cuda.is_available()

import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X / gfx950) and the built HIP library")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def gwn():
    """The HIP library's Python mirror; GPU tests only."""
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test collected on a host without a ROCm device")
    from smallgwn_amd import gwn as g

    return g


@pytest.fixture(scope="session")
def dev():
    import torch

    return torch.device("cuda:0")

"""Shared test configuration: the `gpu` marker and path set-up."""
import os
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
if str(REPO) not in sys.path:
    sys.path.insert(0, str(REPO))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU and liborx_hip.so")
    config.addinivalue_line("markers", "slow: long-running case")


@pytest.fixture(scope="session")
def device():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    from orthoroute_amd import _lib
    _lib.init_device()
    return torch.device("cuda:0")

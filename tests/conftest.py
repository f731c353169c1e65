import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and libgtrap_hip.so")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def gpu():
    """Skip-free GPU fixture: a gpu-marked test fails loudly when no GPU is present."""
    import torch
    from gtrap_amd import _native
    _native.require_gpu()
    torch.cuda.init()
    return torch.device("cuda:0")

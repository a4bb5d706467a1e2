import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle as orc  # oracle/oracle.py
    orc.lib()
    return orc


@pytest.fixture(scope="session")
def hip_device():
    """Skip unless a HIP device is visible (GPU tests are selected with -m gpu)."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    torch.cuda.init()
    return torch.device("cuda:0")

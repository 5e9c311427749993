"""Shared pytest configuration.

Markers: ``gpu`` — needs a HIP device (run on the MI355X box).  Everything else
runs on the CPU: the oracle against closed forms / brute force / golden vectors,
the host logic, and the C-ABI library's load + export + error paths.
"""
import os
import sys
from pathlib import Path

import numpy as np
import pytest

REPO = Path(__file__).resolve().parents[1]
if str(REPO) not in sys.path:
    sys.path.insert(0, str(REPO))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: test needs a HIP GPU (MI355X)")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")


@pytest.fixture
def rng():
    return np.random.default_rng(12345)

"""Shared pytest configuration.

Tests marked ``gpu`` need an MI355X (ROCm device) and the built
``magnelio_amd/libmgh.so``; everything else runs on the CPU.
"""

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built libmgh.so")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def gpu():
    """The torch device of the GPU tests; fails (never skips) without one."""
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test requested but no ROCm device is visible")
    from magnelio_amd import _lib

    _lib.load()
    return torch.device("cuda:0")

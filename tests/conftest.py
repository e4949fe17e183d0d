from __future__ import annotations

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running test")


@pytest.fixture(scope="session")
def gpu():
    """Skip unless a GPU is present; load the HIP library (fails loudly if it is missing)."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from d2ic_amd import _lib

    _lib.load()
    return torch

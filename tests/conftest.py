"""Shared test setup: seeded RNGs and the `gpu` marker (registered in pytest.ini)."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


@pytest.fixture(autouse=True)
def _seed():
    """Reseed NumPy and torch before every test (the reference's conftest.py:17-27 does too)."""
    np.random.seed(0)
    try:
        import torch

        torch.manual_seed(0)
    except ImportError:  # pragma: no cover
        pass
    yield


def gpu_available() -> bool:
    import torch

    return torch.cuda.is_available()


@pytest.fixture
def cuda():
    """Skip-free GPU fixture: a GPU test run without a GPU is an error, not a skip."""
    import torch

    if not torch.cuda.is_available():
        pytest.fail("this test needs a GPU (run with -m 'not gpu' on CPU-only machines)")
    from structsplat_amd import _native

    _native.load()
    return torch.device("cuda")

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "wave-multiphase_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X); run with -m gpu")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def rel_l1(a, b):
    import numpy as np

    den = float(np.abs(b).sum())
    return float(np.abs(a - b).sum()) / (den if den > 0 else 1.0)


@pytest.fixture(scope="session")
def gpu_ok():
    import torch

    # GPU tests run only where a GPU is expected: fail loudly instead of skipping
    assert torch.cuda.is_available(), "GPU test selected but torch.cuda.is_available() is False"
    from wamp import _lib

    _lib.load()
    return True

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.lib()
    return O


def _init_torch_gpu_first():
    """torch ships its own HIP runtime, which must initialise before
    libmagtense_hip's when both drive the device in one process."""
    try:
        import torch
    except ImportError:
        return
    try:
        if torch.cuda.is_available():
            torch.zeros(1, device="cuda")
    except Exception:  # pragma: no cover - no usable device
        pass


_init_torch_gpu_first()

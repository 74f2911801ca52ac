import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (HIP device)")


def _has_gpu() -> bool:
    try:
        import torch

        return bool(torch.cuda.is_available())
    except Exception:  # pragma: no cover
        return False


@pytest.fixture(scope="session")
def gpu():
    """Device handle for GPU tests; errors (does not skip) when no GPU exists,
    so a GPU run without a device is a visible failure."""
    import torch

    if not _has_gpu():
        pytest.fail("GPU test run without a HIP device")
    return torch.device("cuda:0")


@pytest.fixture
def highp():
    """complex128 for the duration of one test (reference tests/conftest.py:62-67)."""
    import tensorcircuit_amd as tc

    tc.set_dtype("complex128")
    yield
    tc.set_dtype("complex64")

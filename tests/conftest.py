import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running")


def gpu_available():
    try:
        from cupyram_amd import _abi
        return _abi.device_count() > 0
    except Exception:
        return False


def _start_torch_first():
    """torch ships its own HIP runtime under the soname of ROCm's
    (libamdhip64.so.7).  Whichever loads first serves the whole process, and
    torch cannot start on ROCm's copy; so torch's goes first, and tests that
    use torch (device buffers, the surrogate) run in any order."""
    try:
        import torch
    except ImportError:
        return
    if torch.cuda.is_available():
        torch.zeros(1, device="cuda")


@pytest.fixture(scope="session")
def require_gpu():
    _start_torch_first()
    from cupyram_amd import _abi
    n = _abi.device_count()
    if n < 1:
        pytest.fail("no HIP device visible; the GPU tests need an MI355X (gfx950)")
    return n

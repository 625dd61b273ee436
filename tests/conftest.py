import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
PKG = REPO / "vectorized-eigensolver_amd"
for p in (str(PKG), str(REPO / "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def _gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Skip-free gate: GPU tests fail loudly when no GPU or no library."""
    import torch
    import veig
    assert torch.cuda.is_available(), "GPU test collected on a host without a GPU"
    veig.lib()
    return torch.device("cuda:0")

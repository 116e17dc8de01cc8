import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "alias-table-gpu_amd")
for p in (PKG, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libatg.so on cuda:0)")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def oracle():
    import oracle_binding
    return oracle_binding


@pytest.fixture(scope="session")
def gpu():
    """Fails loudly if the HIP path is not available on a GPU test."""
    import torch
    import atg
    if not torch.cuda.is_available():
        pytest.fail("GPU test without a visible GPU")
    atg.lib()
    torch.empty(1, device="cuda")
    return torch

import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (REPO, os.path.join(REPO, "tests"), os.path.join(REPO, "x86-simd-sort_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def xss():
    """The HIP library binding; GPU tests fail loudly if it is missing."""
    import torch
    assert torch.cuda.is_available(), "GPU test without a GPU"
    import xss_amd
    xss_amd.lib()
    return xss_amd

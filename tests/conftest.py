import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


@pytest.fixture(scope="session")
def gpu():
    """Skip-free GPU fixture: a gpu-marked test on a machine without a GPU fails
    loudly instead of silently falling back."""
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test started without a visible GPU")
    import paralanczos_amd

    paralanczos_amd.load()
    return torch.device("cuda:0")

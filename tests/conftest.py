import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: full-size cases (seconds to minutes)")


@pytest.fixture(scope="session")
def gpu():
    """Skip-free GPU guard: a gpu-marked test fails loudly without a GPU."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    from rqutils_amd import _lib
    _lib.lib()
    return torch.device("cuda", 0)

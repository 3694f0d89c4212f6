import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built HIP library")


@pytest.fixture(scope="session")
def gpu():
    """cuda:0 with libpowspec_amd.so loaded; GPU tests fail (not skip) if the library is missing."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no ROCm GPU visible")
    from jax_powspec_amd import _lib
    _lib.load()
    return torch.device("cuda", 0)

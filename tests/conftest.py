import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI)")


def _ensure_built():
    import oracle
    from sfft_amd import _lib
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build(with_ref=False)
    if not os.path.exists(_lib.LIB_PATH):
        from sfft_amd import build
        build.build()


_ensure_built()


@pytest.fixture(scope="session")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch

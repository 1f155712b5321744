import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libfem_gpu.so")


def pytest_collection_modifyitems(session, config, items):
    """GPU runs: PyTorch's own HIP runtime has to start before the library's system ROCm runtime in one
    process (the adjoint tests use both), whatever order the tests run in."""
    if any(item.get_closest_marker("gpu") for item in items):
        try:
            import torch
        except ImportError:
            return
        if torch.cuda.is_available():
            torch.cuda.init()


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    oracle.lib()
    return oracle

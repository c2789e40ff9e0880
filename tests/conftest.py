import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X)")


@pytest.fixture(scope="session", autouse=True)
def _torch_runtime_first(request):
    """When GPU tests run, initialise torch's ROCm runtime before the
    library's: torch carries its own HIP runtime, and it finds no device if
    another HIP runtime in the process already opened the GPU."""
    if any(item.get_closest_marker("gpu") for item in request.session.items):
        import torch
        torch.cuda.is_available()
    yield


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle as O
    O.lib()
    return O


def rel_l2(a, b):
    import numpy as np
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


def pytest_collection_finish(session):
    """Initialise torch's device runtime before any test runs when GPU tests are selected,
    so that it comes up first in the process, whatever order the tests run in (the engine
    library brings up the HIP runtime it links when its first handle is created)."""
    if any(item.get_closest_marker("gpu") for item in session.items):
        gpu_available()


@pytest.fixture(autouse=True, scope="module")
def _release_torch_cache():
    """After each test module, return what torch's caching allocator holds to the device:
    the differentiable tests keep trajectories of up to ~200 GB, and the engine's own
    allocations in later modules (hipMalloc, outside torch) need that memory."""
    yield
    torch = sys.modules.get("torch")
    if torch is not None and torch.cuda.is_initialized():
        torch.cuda.empty_cache()


@pytest.fixture(scope="session")
def golden():
    return np.load(os.path.join(ROOT, "tests", "golden", "phenotype_golden.npz"))


@pytest.fixture(scope="session")
def golden_grad():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "grad_fd_golden.json")) as f:
        return json.load(f)


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def require_gpu():
    if not gpu_available():
        pytest.fail("GPU test collected without a visible GPU")

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def to_dev(a, torch):
    """Fortran (nx,ny,nz) numpy array -> contiguous CUDA tensor of shape (nz,ny,nx)."""
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float64).T)).cuda()


def to_host(t):
    """CUDA tensor (nz,ny,nx) -> Fortran (nx,ny,nz) numpy array."""
    return np.asfortranarray(t.detach().cpu().numpy().T)


@pytest.fixture(scope="session")
def torch_gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    return torch

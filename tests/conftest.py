import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")


@pytest.fixture(scope="session")
def gpu():
    """The HIP device as torch sees it; the library must be built -- a missing
    library on a machine with a GPU is a failure, not a skip."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import apxchol_amd
    apxchol_amd.load_library()
    return torch.device("cuda:0")


@pytest.fixture(scope="session")
def rng():
    return np.random.default_rng(12345)


def pytest_collection_modifyitems(config, items):
    # Keep GPU tests in one process and in file order (no xdist reshuffling).
    pass

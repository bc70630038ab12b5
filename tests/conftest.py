import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
if str(ROOT / "tests") not in sys.path:
    sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and runs the HIP library")


@pytest.fixture(scope="session")
def gpu():
    """Device fixture for GPU tests: fails (does not skip) when no GPU is visible, because a GPU
    test that cannot reach the HIP path must not pass."""
    import torch

    assert torch.cuda.is_available(), "GPU test selected but no GPU is visible"
    return torch.device("cuda:0")


def rel_l2(a, b) -> float:
    import numpy as np

    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a visible MI355X GPU and the built HIP library")


def random_model(n: int, U: float, seed: int):
    """Seeded random spin model with diagonal loss/gain (test workload)."""

    from c3lindblad_amd.model import FermionicLindbladModel

    rng = np.random.default_rng(seed)
    E = rng.normal(size=(n, n))
    E = 0.5 * (E + E.T)
    return FermionicLindbladModel(
        E=E,
        gamma_loss=np.diag(rng.uniform(0.05, 0.5, n)),
        gamma_gain=np.diag(rng.uniform(0.05, 0.5, n)),
        U=U,
    )


@pytest.fixture(scope="session")
def torch_gpu():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    return torch

"""Seeded synthetic models of the benchmark workloads.

* ``readme_model``: the reference README quick-start model (3 sites, U=1.2 on
  site 0, ``README.md:62-67``), charge +1, 81 frequencies on [-4, 4], eta 0.
* ``siam_model``: single-impurity Anderson model with Lindblad pseudomodes:
  impurity level -U/2, ``modes`` energies uniform on ``[-span, span]``,
  impurity-mode hoppings ``N(0, hop²)`` and per-mode loss/gain uniform on
  ``gamma_range`` with loss > gain above zero energy and gain > loss below;
  the impurity itself has no loss or gain.  Seeds are
  ``20261203 + 1000*config + instance``.
"""

from __future__ import annotations

import numpy as np

from .model import FermionicLindbladModel

SEED_BASE = 20261203


def readme_model() -> FermionicLindbladModel:
    return FermionicLindbladModel(
        E=np.array([[-0.6, 0.28, 0.28], [0.28, -1.2, 0.0], [0.28, 0.0, 1.2]]),
        gamma_loss=np.diag([0.0, 0.06, 0.48]),
        gamma_gain=np.diag([0.0, 0.48, 0.06]),
        U=1.2,
    )


def readme_grid() -> np.ndarray:
    return np.linspace(-4.0, 4.0, 81)


def siam_model(
    modes: int,
    U: float,
    *,
    seed: int,
    epsilon_d: float | None = None,
    span: float = 4.0,
    hop: float = 0.15,
    gamma_range: tuple[float, float] = (0.02, 0.5),
) -> FermionicLindbladModel:
    rng = np.random.default_rng(seed)
    n = modes + 1
    E = np.zeros((n, n))
    E[0, 0] = -0.5 * U if epsilon_d is None else epsilon_d
    energies = rng.uniform(-span, span, modes)
    hoppings = rng.normal(0.0, hop, modes)
    E[np.arange(1, n), np.arange(1, n)] = energies
    E[0, 1:] = hoppings
    E[1:, 0] = hoppings
    a = rng.uniform(*gamma_range, modes)
    b = rng.uniform(*gamma_range, modes)
    big, small = np.maximum(a, b), np.minimum(a, b)
    loss = np.where(energies > 0.0, big, small)
    gain = np.where(energies > 0.0, small, big)
    return FermionicLindbladModel(
        E=E,
        gamma_loss=np.diag(np.concatenate(([0.0], loss))),
        gamma_gain=np.diag(np.concatenate(([0.0], gain))),
        U=U,
    )


def config2_model(instance: int = 0) -> FermionicLindbladModel:
    """24 pseudomodes, U = 4, eps_d = -2 (single SIAM of the FP64-vs-mixed workload)."""

    return siam_model(24, 4.0, seed=SEED_BASE + 1000 * 2 + instance)


def config2_grid() -> np.ndarray:
    return np.linspace(-8.0, 8.0, 4001)


def config3_model(instance: int) -> FermionicLindbladModel:
    """One instance of the uneven-size SIAM batch (4..32 modes, U in [0, 6])."""

    rng = np.random.default_rng(SEED_BASE + 1000 * 3 + instance)
    modes = int(rng.integers(4, 33))
    U = float(rng.uniform(0.0, 6.0))
    eps = -0.5 * U + float(rng.normal(0.0, 0.5))
    return siam_model(modes, U, seed=SEED_BASE + 1000 * 3 + 100000 + instance, epsilon_d=eps, span=5.0, hop=0.2,
                      gamma_range=(0.01, 0.6))

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


def config3_modes(instance: int) -> int:
    """Pseudomode count of one instance of the uneven-size batch (uniform on 4..32)."""

    return int(np.random.default_rng(SEED_BASE + 1000 * 3 + instance).integers(4, 33))


def config3_model(instance: int) -> FermionicLindbladModel:
    """One instance of the uneven-size SIAM batch (4..32 modes, U in [0, 6])."""

    rng = np.random.default_rng(SEED_BASE + 1000 * 3 + instance)
    modes = int(rng.integers(4, 33))
    U = float(rng.uniform(0.0, 6.0))
    eps = -0.5 * U + float(rng.normal(0.0, 0.5))
    return siam_model(modes, U, seed=SEED_BASE + 1000 * 3 + 100000 + instance, epsilon_d=eps, span=5.0, hop=0.2,
                      gamma_range=(0.01, 0.6))


CONFIG3_INSTANCES = 512


def config3_grid() -> np.ndarray:
    """The batch's 2048 frequencies per instance."""

    return np.linspace(-8.0, 8.0, 2048)


def _symmetric_orbit_values(orbitals: int, rng, scale: float) -> np.ndarray:
    """Real ``V[p,q,r,s]`` with the 8-fold permutational symmetry of real orbitals
    (``(pq|rs) = (qp|rs) = (pq|sr) = (rs|pq) = ...``): one ``N(0, scale^2)`` value per orbit of
    index tuples, drawn in lexicographic order of the orbit's smallest member."""

    def orbit(p, q, r, s):
        return {(p, q, r, s), (q, p, r, s), (p, q, s, r), (q, p, s, r),
                (r, s, p, q), (s, r, p, q), (r, s, q, p), (s, r, q, p)}

    V = np.zeros((orbitals,) * 4)
    seen = set()
    for index in np.ndindex(*V.shape):
        if index in seen:
            continue
        members = orbit(*index)
        value = float(rng.normal(0.0, scale))
        for m in members:
            V[m] = value
        seen |= members
    return V


def multiorbital_model(
    orbitals: int,
    modes_per_orbital,
    *,
    seed: int,
    U: float = 3.0,
    J: float = 0.5,
    perturbation: float = 0.1,
    span: float = 6.0,
    hop: float = 0.2,
    gamma_range: tuple[float, float] = (0.02, 0.5),
    bias: float = 0.5,
    epsilon: float | None = None,
    impurity_hopping: float = 0.0,
) -> FermionicLindbladModel:
    """Multi-orbital impurity with a general two-body interaction and Lindblad pseudomodes.

    Interaction (spatial, chemists' order, ``H = 1/2 sum (ab|cd) c†_{aσ} c†_{cσ'} c_{dσ'} c_{bσ}``):
    Kanamori ``(aa|aa) = U``, ``(aa|bb) = U' = U - 2J``, ``(ab|ba) = (ab|ab) = J`` (so opposite-spin
    density ``U - 2J``, same-spin ``U - 3J``, spin flip and pair hopping ``J``) plus a random
    real tensor with the 8-fold symmetry, ``N(0, perturbation^2)`` per symmetry orbit.  It enters
    the model as a ``two_body_interaction`` on the impurity orbitals ``0..orbitals-1`` through the
    spatial-to-spin-orbital map of ``interaction_from_spatial_tensor``
    (``cc/_static_schema.py:77-91``).  Impurity levels sit at the half-filling point
    ``-(U + 2(K-1)U' - (K-1)J)/2`` unless ``epsilon`` is given.  Orbital ``a`` couples to its own
    ``modes_per_orbital[a]`` pseudomodes (an int applies to every orbital): energies uniform on
    ``[-span, span]``, hoppings ``N(0, hop^2)``, loss and gain uniform on ``gamma_range``; the first
    half of each orbital's modes belongs to a left lead at chemical potential ``+bias`` and the
    rest to a right lead at ``-bias``, the larger rate being the loss above the lead's potential
    and the gain below it.
    """

    from .model import SparseTwoBodyInteraction

    rng = np.random.default_rng(seed)
    K = int(orbitals)
    counts = [int(modes_per_orbital)] * K if np.isscalar(modes_per_orbital) else [int(m) for m in modes_per_orbital]
    if len(counts) != K or K < 1 or min(counts) < 0:
        raise ValueError("modes_per_orbital must give a non-negative count per orbital")
    n = K + sum(counts)
    Up = U - 2.0 * J
    level = -(U + 2 * (K - 1) * Up - (K - 1) * J) / 2.0 if epsilon is None else float(epsilon)
    E = np.zeros((n, n))
    loss = np.zeros(n)
    gain = np.zeros(n)
    for a in range(K):
        E[a, a] = level
        for b in range(a + 1, K):
            E[a, b] = E[b, a] = impurity_hopping
    mode = K
    for a, count in enumerate(counts):
        for k in range(count):
            mu = bias if k < (count + 1) // 2 else -bias
            E[mode, mode] = rng.uniform(-span, span)
            E[a, mode] = E[mode, a] = rng.normal(0.0, hop)
            x, y = rng.uniform(*gamma_range, 2)
            big, small = max(x, y), min(x, y)
            loss[mode], gain[mode] = (big, small) if E[mode, mode] > mu else (small, big)
            mode += 1
    C = np.zeros((K,) * 4)
    for a in range(K):
        C[a, a, a, a] = U
        for b in range(K):
            if a != b:
                C[a, a, b, b] = Up
                C[a, b, b, a] = J
                C[a, b, a, b] = J
    C = C + _symmetric_orbit_values(K, rng, perturbation)
    terms = []
    for w in range(K):
        for x in range(K):
            for y in range(K):
                for z in range(K):
                    value = 0.5 * C[w, z, x, y]  # T[w,x,y,z] of interaction_from_spatial_tensor
                    if value == 0.0:
                        continue
                    for s1 in (0, 1):
                        for s2 in (0, 1):
                            creation = (2 * w + s1, 2 * x + s2)
                            annihilation = (2 * z + s1, 2 * y + s2)
                            if creation[0] != creation[1] and annihilation[0] != annihilation[1]:
                                terms.append((*creation, *annihilation, complex(value)))
    return FermionicLindbladModel(
        E=E,
        gamma_loss=np.diag(loss),
        gamma_gain=np.diag(gain),
        U=0.0,
        interaction_sites=tuple(range(K)),
        two_body_interaction=SparseTwoBodyInteraction(2 * n, tuple(terms)),
    )


def config5_model(instance: int = 0) -> FermionicLindbladModel:
    """Three Kanamori orbitals (U = 3, J = 0.5, U' = 2) + V_pqrs ~ N(0, 0.1^2), 12 pseudomodes each
    (n = 39), lead bias +-0.5."""

    return multiorbital_model(3, 12, seed=SEED_BASE + 1000 * 5 + instance)


def config5_grid() -> np.ndarray:
    return np.linspace(-8.0, 8.0, 8001)


def s135_validation_model() -> FermionicLindbladModel:
    """The compact interacting impurity of the reference's full-S135 tests
    (``tests/s135_fixture.py:8-28``): impurity level -U/2 with U = 1.2, two bath sites at -1.2 and
    1.2 coupled by 0.28 + 0.04 (1 - |e| / max|e|), losses 0.06 + 0.42 s and gains 0.48 - 0.42 s on
    the bath (s the bath level scaled to [0, 1])."""

    U = 1.2
    bath = np.linspace(-1.2, 1.2, 2)
    E = np.diag(np.concatenate(([-U / 2.0], bath))).astype(np.complex128)
    coupling = 0.28 + 0.04 * (1.0 - np.abs(bath) / np.max(np.abs(bath)))
    E[0, 1:] = coupling
    E[1:, 0] = coupling
    scaled = (bath - bath[0]) / (bath[-1] - bath[0])
    loss = np.zeros(3)
    gain = np.zeros(3)
    loss[1:] = 0.06 + 0.42 * scaled
    gain[1:] = 0.48 - 0.42 * scaled
    return FermionicLindbladModel(E, np.diag(loss), np.diag(gain), U=U)

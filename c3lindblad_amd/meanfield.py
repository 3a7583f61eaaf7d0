"""Gaussian (Hartree--Fock) mean field of a fermionic Lindblad model.

Restates ``src/c3l/quadratic_ness.py:20-50`` (quadratic steady covariance),
``src/c3l/gaussian_mean_field.py:199-254`` (sparse Wick/Fock contraction) and
the unrestricted branch of ``_solve_mean_field``/``solve_ghf``
(``gaussian_mean_field.py:390-714``): damped Pulay (DIIS) mixing of the fixed
point ``gamma = C[F[gamma]]`` with the same defaults, convergence test and
final trial step.  Host-side set-up only; it is not on the device path.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.linalg

from .model import SparseTwoBodyInteraction


def finite_pm_covariance(E, gamma_loss, gamma_gain) -> tuple[np.ndarray, float]:
    """Solve ``A g + g A^† + 2 Γ_gain = 0`` with ``A = -iE - Γ_loss - Γ_gain``.

    ``g[p,q] = <a_q^† a_p>``; returns ``(g, relative Lyapunov residual)``
    (``quadratic_ness.py:20-50``).
    """

    h = np.asarray(E, dtype=np.complex128)
    loss = np.asarray(gamma_loss, dtype=np.complex128)
    gain = np.asarray(gamma_gain, dtype=np.complex128)
    A = -1j * h - loss - gain
    g = scipy.linalg.solve_sylvester(A, A.conj().T, -2.0 * gain)
    R = A @ g + g @ A.conj().T + 2.0 * gain
    scale = max(np.linalg.norm(A, ord=2) * np.linalg.norm(g), np.linalg.norm(2.0 * gain), 1.0)
    return g, float(np.linalg.norm(R) / scale)


def wick_fock(h, interaction: SparseTwoBodyInteraction, gamma) -> np.ndarray:
    """``F_pq = h_pq + sum_rs vbar_prqs gamma_sr`` over canonical sparse terms.

    The four signed updates per stored term are the antisymmetric images of
    one canonical entry (``gaussian_mean_field.py:241-248``).
    """

    out = np.array(h, dtype=np.complex128, copy=True)
    d = np.asarray(gamma, dtype=np.complex128)
    for p, r, q, s, v in interaction.terms:
        out[p, q] += v * d[s, r]
        out[r, q] -= v * d[s, p]
        out[p, s] -= v * d[q, r]
        out[r, s] += v * d[q, p]
    return out


@dataclass(frozen=True)
class GHFResult:
    converged: bool
    gamma: np.ndarray
    fock: np.ndarray
    iterations: int
    self_consistency_residual: float
    lyapunov_residual: float
    physicality_defect: float


def _lyapunov_residual(F, loss, gain, gamma) -> float:
    A = -1j * F - loss - gain
    R = A @ gamma + gamma @ A.conj().T + 2.0 * gain
    scale = max(np.linalg.norm(A, ord=2) * np.linalg.norm(gamma), np.linalg.norm(2.0 * gain), 1.0)
    return float(np.linalg.norm(R) / scale)


def solve_ghf(
    h_bare,
    interaction: SparseTwoBodyInteraction,
    gamma_loss,
    gamma_gain,
    *,
    tolerance: float = 1.0e-10,
    max_iterations: int = 200,
    damping: float = 0.5,
    diis_start: int = 3,
    diis_size: int = 6,
) -> GHFResult:
    """Unrestricted full-HF fixed point by damped deterministic Pulay mixing."""

    h = np.asarray(h_bare, dtype=np.complex128)
    loss = np.asarray(gamma_loss, dtype=np.complex128)
    gain = np.asarray(gamma_gain, dtype=np.complex128)
    if interaction.n_spin_orbitals != h.shape[0]:
        raise ValueError("the sparse interaction does not match h_bare")

    def fock(density):
        return wick_fock(h, interaction, density)

    gamma, _ = finite_pm_covariance(h, loss, gain)
    targets: list[np.ndarray] = []
    residuals: list[np.ndarray] = []
    converged = False
    F = fock(gamma)
    target = gamma
    iterations = 0
    for iteration in range(max_iterations):
        iterations = iteration + 1
        F = fock(gamma)
        target, _ = finite_pm_covariance(F, loss, gain)
        residual = target - gamma
        residual_norm = float(np.linalg.norm(residual) / max(np.linalg.norm(target), 1.0))
        targets.append(target)
        residuals.append(residual)
        if len(targets) > diis_size:
            targets.pop(0)
            residuals.pop(0)
        if iteration + 1 >= diis_start and len(targets) > 1:
            m = len(targets)
            B = np.empty((m + 1, m + 1), dtype=np.float64)
            for i in range(m):
                for j in range(m):
                    B[i, j] = np.vdot(residuals[i], residuals[j]).real
            B[:m, m] = 1.0
            B[m, :m] = 1.0
            B[m, m] = 0.0
            rhs = np.zeros(m + 1)
            rhs[m] = 1.0
            c = np.linalg.lstsq(B, rhs, rcond=None)[0][:m]
            mixed = sum(ci * ti for ci, ti in zip(c, targets))
        else:
            mixed = target
        update = damping * (mixed - gamma)
        if residual_norm <= tolerance:
            trial_gamma = np.asarray(target)
            trial_fock = fock(trial_gamma)
            trial_target, _ = finite_pm_covariance(trial_fock, loss, gain)
            trial = float(np.linalg.norm(trial_target - trial_gamma) / max(np.linalg.norm(trial_target), 1.0))
            if trial <= tolerance:
                gamma, F, target = trial_gamma, trial_fock, trial_target
                converged = True
                break
        gamma = gamma + update
    if not converged:
        F = fock(gamma)
        target, _ = finite_pm_covariance(F, loss, gain)
    sc = float(np.linalg.norm(target - gamma) / max(np.linalg.norm(target), 1.0))
    w = np.linalg.eigvalsh(0.5 * (gamma + gamma.conj().T))
    physicality = max(0.0, float(-w[0]), float(w[-1] - 1.0))
    return GHFResult(converged, gamma, F, iterations, sc, _lyapunov_residual(F, loss, gain, gamma), physicality)


def quadratic_retarded_green(F, gamma_loss, gamma_gain, z) -> np.ndarray:
    """``G^R(z) = [z I - F + i(Γ_loss+Γ_gain)]^{-1}`` (``gaussian_mean_field.py:827-872``)."""

    F = np.asarray(F, dtype=np.complex128)
    A = 1j * (np.asarray(gamma_loss) + np.asarray(gamma_gain)) - F
    eye = np.eye(F.shape[0], dtype=np.complex128)
    return np.stack([np.linalg.inv(complex(s) * eye + A) for s in np.atleast_1d(z)])

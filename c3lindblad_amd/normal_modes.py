"""Quadratic superfermion normal modes and their anti-linear tilde pairing.

Restates, on the host in FP64:

* ``quadratic_mode_matrix`` (``src/c3l/superfermion/generator.py:10-22``);
* ``normal_mode_reference_from_arrays`` (``superfermion/normal_modes.py:484-618``):
  eigen-decomposition of ``h``, unit-norm columns with the largest-magnitude
  pivot made real positive, the tilde-Lagrangian neutrality projection when
  the computed half-plane subspaces are not exactly exchange-neutral
  (``normal_modes.py:260-452``), the frozen ordering (negative imaginary part
  first, then by imaginary and real part), ``V_inv``, vacuum mask, mode
  charges and rapidities;
* ``tilde_map_from_references`` (``ci/tilde.py:133-251``): the monomial
  tilde action ``tilde(bar P_p) = factor_p bar P_partner(p)`` found by a
  least-squares fit plus linear assignment.

Strictly degenerate spectra (``normal_modes.py:118-204``) are rejected with
an error instead of being re-based; none of the benchmark models hit them.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.linalg
from scipy.optimize import linear_sum_assignment

from .model import TolerancePolicy


def quadratic_mode_matrix(E, gamma_loss, gamma_gain) -> np.ndarray:
    """``h = [[E + iΩ, 2Γ_gain], [-2Γ_loss, E - iΩ]]`` with ``Ω = Γ_gain - Γ_loss``."""

    E = np.asarray(E, dtype=np.complex128)
    loss = np.asarray(gamma_loss, dtype=np.complex128)
    gain = np.asarray(gamma_gain, dtype=np.complex128)
    omega = gain - loss
    return np.block([[E + 1j * omega, 2.0 * gain], [-2.0 * loss, E - 1j * omega]]).astype(np.complex128)


def _mode_pivot(vector: np.ndarray) -> int:
    mags = np.abs(vector)
    top = float(np.max(mags))
    tie = 128.0 * np.finfo(np.float64).eps * max(top, 1.0) * vector.size
    return int(np.flatnonzero(mags >= top - tie)[0])


def _normalize_mode_columns(vectors: np.ndarray) -> None:
    for p in range(vectors.shape[1]):
        norm = np.linalg.norm(vectors[:, p])
        if norm == 0.0:
            raise ValueError("normal-mode eigensolver returned a zero eigenvector")
        vectors[:, p] /= norm
        pivot = _mode_pivot(vectors[:, p])
        vectors[:, p] *= np.exp(-1j * np.angle(vectors[pivot, p]))
        if vectors[pivot, p].real < 0.0:
            vectors[:, p] *= -1.0


def _exchange(n: int) -> np.ndarray:
    eye = np.eye(n, dtype=np.complex128)
    zero = np.zeros((n, n), dtype=np.complex128)
    return np.block([[zero, eye], [eye, zero]])


def _neutral_graph(block: np.ndarray, transform: np.ndarray, n: int, limit: float) -> np.ndarray:
    """Replace the graph ``[I; R]`` of a half-plane subspace by its polar-unitary graph."""

    coords = transform.conj().T @ block
    upper, lower = coords[:n], coords[n:]
    cond = float(np.linalg.cond(upper, p=2))
    if not np.isfinite(cond) or cond * np.finfo(np.float64).eps > limit:
        raise ValueError("tilde-Lagrangian graph chart is ill conditioned")
    graph = scipy.linalg.solve(upper.T, lower.T, assume_a="gen").T
    left, _, right = scipy.linalg.svd(graph, full_matrices=False)
    return transform @ np.vstack((upper, (left @ right) @ upper))


def _tilde_lagrangian_projection(h, values, vectors, policy: TolerancePolicy) -> tuple[np.ndarray, bool]:
    """Project numerically non-neutral half-plane subspaces onto exact neutrality."""

    dim = h.shape[0]
    n = dim // 2
    negative = np.flatnonzero(values.imag < 0.0)
    positive = np.flatnonzero(values.imag > 0.0)
    if negative.size != n or positive.size != n:
        return vectors, False
    X = _exchange(n)
    stable = vectors[:, negative]
    pos = vectors[:, positive]
    before = max(
        np.linalg.norm(stable.conj().T @ X @ stable, ord=2),
        np.linalg.norm(pos.conj().T @ X @ pos, ord=2),
    )
    if before <= policy.algebra_abs_rel.absolute:
        return vectors, False
    eye = np.eye(n, dtype=np.complex128)
    transform = np.block([[eye, eye], [eye, -eye]]) / np.sqrt(2.0)
    limit = policy.normal_mode_condition_budget_fraction * policy.algebra_abs_rel.threshold(1.0)
    projected_stable = _neutral_graph(stable, transform, n, limit)
    cost = np.abs(values[negative, None].conj() - values[positive][None, :])
    rows, cols = linear_sum_assignment(cost)
    if not np.array_equal(rows, np.arange(n)):
        raise ValueError("conjugate mode assignment is incomplete")
    pair_threshold = policy.eigensolver_abs_rel.threshold(max(float(np.linalg.norm(h, ord=2)), 1.0))
    if np.max(cost[rows, cols]) > pair_threshold:
        raise ValueError("normal-mode spectrum lacks conjugate tilde partners")
    paired = positive[cols]
    neutral_positive = _neutral_graph(vectors[:, paired], transform, n, limit)
    overlap = (X @ projected_stable.conj()).T @ neutral_positive
    cond = float(np.linalg.cond(overlap, p=2))
    if not np.isfinite(cond) or cond * np.finfo(np.float64).eps > limit:
        raise ValueError("tilde-paired mode overlap is ill conditioned")
    projected_positive = scipy.linalg.solve(overlap.T, neutral_positive.T, assume_a="gen").T
    out = np.array(vectors, dtype=np.complex128, copy=True)
    out[:, negative] = projected_stable
    out[:, paired] = projected_positive
    _normalize_mode_columns(out)
    after = np.linalg.norm(out[:, negative].conj().T @ X @ out[:, negative], ord=2)
    if after > policy.algebra_abs_rel.threshold(1.0):
        raise ValueError("tilde-Lagrangian projection did not restore neutrality")
    return out, True


@dataclass(frozen=True)
class NormalModeReference:
    """Biorthogonal one-spin modes: ``h V = V diag(eigenvalues)``."""

    h: np.ndarray
    eigenvalues: np.ndarray
    V: np.ndarray
    V_inv: np.ndarray
    vacuum_mask: np.ndarray
    mode_charges: np.ndarray
    rapidities: np.ndarray
    canonicalization_applied: bool

    @property
    def n_orbitals(self) -> int:
        return self.h.shape[0] // 2


def normal_mode_reference(E, gamma_loss, gamma_gain, policy: TolerancePolicy | None = None) -> NormalModeReference:
    policy = TolerancePolicy.first() if policy is None else policy
    h = quadratic_mode_matrix(E, gamma_loss, gamma_gain)
    values, V = scipy.linalg.eig(h)
    h_scale = max(float(np.linalg.norm(h, ord=2)), 1.0)
    strict = max(policy.input_abs_rel.absolute, 128.0 * np.finfo(np.float64).eps * h_scale)
    for half in (values.imag < 0.0, values.imag > 0.0):
        sel = values[half]
        gaps = np.abs(sel[:, None] - sel[None, :])
        np.fill_diagonal(gaps, np.inf)
        if sel.size > 1 and np.min(gaps) <= strict:
            raise NotImplementedError("strictly degenerate normal modes are not supported")
    _normalize_mode_columns(V)
    V, applied = _tilde_lagrangian_projection(h, values, V, policy)
    occupation_threshold = policy.eigensolver_abs_rel.threshold(np.linalg.norm(h, ord=2))
    if np.min(np.abs(values.imag)) <= occupation_threshold:
        raise ValueError("normal-mode occupation is ambiguous")
    n = h.shape[0] // 2
    if np.count_nonzero(values.imag < 0.0) != n:
        raise ValueError("normal-mode spectrum must split equally across damping half-planes")
    order = np.array(
        sorted(range(values.size), key=lambda p: (0 if values[p].imag < 0.0 else 1, values[p].imag, values[p].real)),
        dtype=np.int64,
    )
    values = values[order]
    V = V[:, order]
    V_inv = scipy.linalg.inv(V)
    cond = float(np.linalg.cond(V, p=2))
    limit = policy.normal_mode_condition_budget_fraction * policy.algebra_abs_rel.threshold(1.0)
    if cond * np.finfo(np.float64).eps > limit:
        raise ValueError("normal-mode eigenvectors are too ill conditioned")
    vacuum = values.imag > 0.0
    charges = np.where(vacuum, -1, 1).astype(np.int8)
    rapidities = (1 - 2 * vacuum.astype(np.int8)) * values
    loss = np.asarray(gamma_loss, dtype=np.complex128)
    gain = np.asarray(gamma_gain, dtype=np.complex128)
    eta = complex(-np.trace(np.asarray(E)) - 1j * np.trace(loss + gain))
    if abs(eta + np.sum(values[vacuum])) > policy.algebra_abs_rel.threshold(max(abs(eta), 1.0)):
        raise ValueError("occupied-mode scalar cancellation exceeds policy")
    return NormalModeReference(h, values, V, V_inv, vacuum, charges, rapidities, applied)


def _scaled_fit(target: np.ndarray, source: np.ndarray) -> tuple[complex, float]:
    denom = np.vdot(source, source)
    if denom == 0.0:
        raise ValueError("cannot fit a zero tilde source vector")
    alpha = complex(np.vdot(source, target) / denom)
    residual = float(np.linalg.norm(target - alpha * source) / max(float(np.linalg.norm(target)), 1.0e-13))
    return alpha, residual


def _local_tilde_map(ref: NormalModeReference, policy: TolerancePolicy):
    n = ref.n_orbitals
    negative = np.flatnonzero(~ref.vacuum_mask)
    positive = np.flatnonzero(ref.vacuum_mask)
    X = _exchange(n)
    costs = np.empty((n, n))
    fits = {}
    for row, p in enumerate(negative):
        source = X @ ref.V[:, p].conj()
        for col, q in enumerate(positive):
            fits[row, col] = _scaled_fit(ref.V_inv[q, :], source)
            costs[row, col] = fits[row, col][1]
    rows, cols = linear_sum_assignment(costs)
    threshold = policy.eigensolver_abs_rel.threshold(1.0)
    if np.max(costs[rows, cols]) > threshold:
        raise ValueError("tilde action is not monomial in the accepted normal modes")
    partners = np.empty(2 * n, dtype=np.int64)
    factors = np.empty(2 * n, dtype=np.complex128)
    residuals = np.empty(2 * n)
    for row, col in zip(rows, cols):
        p, q = int(negative[row]), int(positive[col])
        alpha, res = fits[row, col]
        partners[p], factors[p] = q, 1.0 / alpha
        beta, rev = _scaled_fit(ref.V[:, p], X @ ref.V_inv[q, :].conj())
        partners[q], factors[q] = p, 1.0 / beta
        cq, cq_res = _scaled_fit(ref.V_inv[p, :], X @ ref.V[:, q].conj())
        cp, cp_res = _scaled_fit(ref.V[:, q], X @ ref.V_inv[p, :].conj())
        dq = abs(cq - factors[q]) / max(abs(cq), abs(factors[q]), 1.0e-13)
        dp = abs(cp - factors[p]) / max(abs(cp), abs(factors[p]), 1.0e-13)
        residuals[p] = max(res, cp_res, dp)
        residuals[q] = max(rev, cq_res, dq)
    return partners, factors, residuals


@dataclass(frozen=True)
class ModeTilde:
    """Mode-level tilde map over all ``4n`` particle-hole modes (spin order up, down)."""

    partners: np.ndarray
    factors: np.ndarray


def tilde_map_from_references(up: NormalModeReference, down: NormalModeReference, policy=None) -> ModeTilde:
    policy = TolerancePolicy.first() if policy is None else policy
    n = up.n_orbitals
    partners = np.empty(4 * n, dtype=np.int64)
    factors = np.empty(4 * n, dtype=np.complex128)
    residuals = np.empty(4 * n)
    for spin, ref in enumerate((up, down)):
        p, f, r = _local_tilde_map(ref, policy)
        base = 2 * n * spin
        partners[base : base + 2 * n] = p + base
        factors[base : base + 2 * n] = f
        residuals[base : base + 2 * n] = r
    if np.max(residuals) > policy.eigensolver_abs_rel.threshold(1.0):
        raise ValueError("derived tilde directional residual exceeds policy")
    if np.max(np.abs(factors.conj() * factors[partners] - 1.0)) > policy.algebra_abs_rel.threshold(1.0):
        raise ValueError("derived mode factors violate double-tilde invariance")
    return ModeTilde(partners, factors)

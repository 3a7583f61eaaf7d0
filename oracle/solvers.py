"""FP64 CPU solvers -- TEST INFRASTRUCTURE / CPU BASELINE ONLY.

* ``solve_ccsd``: inexact Newton with SciPy GMRES(restart 40, 12 cycles),
  rapidity preconditioner and Armijo backtracking, legacy forcing -- a
  restatement of ``cc/component_solver.py:103-295`` and
  ``cc/newton_krylov.py:137-237`` over any residual/JVP callables.
* ``dense_matrix``: columns of a small linear action.
* ``retarded_green_dense``: ``G(ω) = sink·(z-σ)^{-1}(add) + sink·(z-σ)^{-1}(rem)``
  by dense solves (``cc/component_response.py:180-245`` at convergence).
* ``arnoldi_resolvent``: two-pass CGS Arnoldi with Schur-based shifted
  solves (``green/krylov.py:284-417``).

Pins: the Newton root is pinned by its residual norm (any root with
``||pack R|| <= 1e-10`` is the reference root up to conditioning); the dense
Green function is pinned against Lindblad ED at n = 1 and against the
closed-form quadratic resolvent at U = 0 (``tests/test_physics_pins.py``).
"""

from __future__ import annotations

import numpy as np
import scipy.linalg
from scipy.sparse.linalg import LinearOperator, gmres


def preconditioner_diagonal(rapidity_sums: np.ndarray) -> np.ndarray:
    d = -1j * np.asarray(rapidity_sums, dtype=np.complex128)
    mags = np.abs(d)
    threshold = 1.0e-12 * max(float(np.max(mags, initial=0.0)), 1.0)
    small = mags < threshold
    d[small] = threshold * np.exp(1j * np.angle(d[small] + 1e-300))
    return d


def solve_ccsd(tilde, rapidity_sums, residual, jvp_at, *, tolerance=1e-10, max_newton=20, restart=40, cycles=12,
               max_line=14, armijo=1e-4, x0=None):
    """Return ``(x, amplitudes, packed_norm, iterations, counts)``.

    ``residual(t) -> R`` (complex) and ``jvp_at(t) -> (dt -> J dt)``.
    """

    D = tilde.real_size
    x = np.zeros(D) if x0 is None else np.array(x0, dtype=np.float64)
    diag = preconditioner_diagonal(rapidity_sums)
    M = LinearOperator((D, D), matvec=lambda v: tilde.pack(tilde.unpack(v) / diag, check=False), dtype=np.float64)
    counts = {"residual": 0, "jvp": 0}

    def real_residual(v):
        counts["residual"] += 1
        return tilde.pack(residual(tilde.unpack(v)), check=False)

    iterations = 0
    for it in range(max_newton + 1):
        r = real_residual(x)
        norm = float(np.linalg.norm(r))
        if norm <= tolerance or it == max_newton:
            break
        J = jvp_at(tilde.unpack(x))

        def mv(v):
            counts["jvp"] += 1
            return tilde.pack(J(tilde.unpack(v)), check=False)

        eta = min(1e-3, max(1e-8, 0.1 * max(norm, tolerance)))
        step, _ = gmres(LinearOperator((D, D), matvec=mv, dtype=np.float64), -r, M=M, rtol=eta, atol=0.0,
                        restart=restart, maxiter=cycles)
        damping = 1.0
        accepted = False
        for _ in range(max_line):
            trial = real_residual(x + damping * step)
            if np.linalg.norm(trial) <= (1.0 - armijo * damping) * norm:
                x = x + damping * step
                accepted = True
                break
            damping *= 0.5
        iterations += 1
        if not accepted:
            break
    t = tilde.unpack(x)
    packed = float(np.linalg.norm(tilde.pack(residual(t), check=False)))
    return x, t, packed, iterations, counts


def dense_matrix(matvec, dim: int) -> np.ndarray:
    A = np.empty((dim, dim), dtype=np.complex128)
    for j in range(dim):
        e = np.zeros(dim, dtype=np.complex128)
        e[j] = 1.0
        A[:, j] = matvec(e)
    return A


def retarded_green_dense(A, addition, removal, sink, omega, eta):
    eye = np.eye(A.shape[0], dtype=np.complex128)
    add = np.empty(len(omega), dtype=np.complex128)
    rem = np.empty_like(add)
    for k, w in enumerate(omega):
        M = (w + 1j * eta) * eye - A
        add[k] = sink @ np.linalg.solve(M, addition)
        rem[k] = sink @ np.linalg.solve(M, removal)
    return add, rem


def arnoldi_resolvent(matvec, b, shifts, *, maximum_dimension, tolerance, minimum_dimension=16, check_interval=8,
                      breakdown=1e-14):
    """Return ``(V, H, beta)`` of a two-pass CGS Arnoldi run to the residual tolerance."""

    n = b.size
    beta = float(np.linalg.norm(b))
    V = np.zeros((n, maximum_dimension + 1), dtype=np.complex128, order="F")
    H = np.zeros((maximum_dimension + 1, maximum_dimension), dtype=np.complex128)
    V[:, 0] = b / beta
    actual = maximum_dimension
    for j in range(maximum_dimension):
        w = matvec(V[:, j])
        for _ in range(2):
            c = V[:, : j + 1].conj().T @ w
            H[: j + 1, j] += c
            w = w - V[:, : j + 1] @ c
        H[j + 1, j] = np.linalg.norm(w)
        actual = j + 1
        if abs(H[j + 1, j]) <= breakdown:
            break
        V[:, j + 1] = w / H[j + 1, j]
        if actual >= minimum_dimension and (actual % check_interval == 0 or actual == maximum_dimension):
            if np.max(reduced_residuals(H[: actual + 1, :actual], beta, shifts)) <= tolerance:
                break
    return V[:, : actual + 1], H[: actual + 1, :actual], beta


def reduced_residuals(H, beta, shifts):
    m = H.shape[1]
    T, Q = scipy.linalg.schur(H[:m, :m], output="complex")
    rhs = Q.conj().T[:, 0] * beta
    out = np.empty(len(shifts))
    for k, z in enumerate(shifts):
        y = Q @ scipy.linalg.solve_triangular(z * np.eye(m) - T, rhs)
        out[k] = abs(H[m, m - 1] * y[-1]) / beta
    return out


def arnoldi_green(V, H, beta, sink, shifts):
    m = H.shape[1]
    T, Q = scipy.linalg.schur(H[:m, :m], output="complex")
    rhs = Q.conj().T[:, 0] * beta
    sv = (sink @ V[:, :m]) @ Q
    return np.array([sv @ scipy.linalg.solve_triangular(z * np.eye(m) - T, rhs) for z in shifts])

"""Fock-space brute force and physical Lindblad ED -- TEST INFRASTRUCTURE ONLY.

Particle-hole Fock space (pins ``oracle/polynomial`` and the compiled plans)
---------------------------------------------------------------------------
Sparse Jordan--Wigner matrices for all ``M = 4n`` PH modes.  The canonical
string ``τ_μ = bar P_p1 ... bar P_pk`` (``p1 < ... < pk``) is built as a
literal matrix product, ``K = sum_p λ_p bar P_p P_p + sum_products c prod F``
as a sparse matrix, and ``exp(±T)`` by its finite nilpotent series.  Then

* ``R_μ = -i [exp(-T) K exp(T) |0>]_μ``   (``tests/cc/test_charge_s13.py:192-232``)
* ``σ_ν = [exp(-T) K exp(T) X|0> - X exp(-T) K exp(T)|0>]_ν``

No sign rule of the polynomial oracle is reused: every sign comes from the
matrix products.

Physical Lindblad ED (pins the whole lowering end to end)
---------------------------------------------------------
Column-major vectorised generator of ``tests/reference/exact_generator.py:79-125``
with loss ``Γ_ij (2 c_j ρ c_i† - {c_i† c_j, ρ})`` and gain
``Γ_ij (2 c_i† ρ c_j - {c_j c_i†, ρ})``; the retarded Green function uses the
fermion-odd generator (jump terms with flipped sign) exactly as
``tests/cc/test_charge_s13.py:44-65,109-122``.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def jw_creators(num_modes: int) -> list[sp.csr_matrix]:
    dim = 1 << num_modes
    states = np.arange(dim, dtype=np.int64)
    out = []
    for p in range(num_modes):
        bit = 1 << p
        free = (states & bit) == 0
        src = states[free]
        lower = src & (bit - 1)
        parity = np.array([bin(int(v)).count("1") & 1 for v in lower], dtype=np.int64)
        data = np.where(parity == 1, -1.0, 1.0)
        out.append(sp.csr_matrix((data, (src | bit, src)), shape=(dim, dim), dtype=np.complex128))
    return out


class PHFockSpace:
    """Operators of a lowered problem on the complete PH-mode Fock space."""

    def __init__(self, problem):
        self.M = problem.num_modes
        if self.M > 16:
            raise ValueError("Fock brute force is limited to 16 modes")
        self.dim = 1 << self.M
        self.cre = jw_creators(self.M)
        self.ann = [c.T.tocsr() for c in self.cre]
        K = sp.csr_matrix((self.dim, self.dim), dtype=np.complex128)
        for p, lam in enumerate(problem.rapidities):
            K = K + lam * (self.cre[p] @ self.ann[p])
        for record in problem.products:
            prod = sp.identity(self.dim, dtype=np.complex128, format="csr")
            for factor in record.factors:
                prod = prod @ self.linear(factor.components)
            K = K + record.coefficient * prod
        self.K = K.tocsr()

    def linear(self, components) -> sp.csr_matrix:
        F = sp.csr_matrix((self.dim, self.dim), dtype=np.complex128)
        for kind, mode, c in components:
            F = F + c * (self.cre[mode] if kind == "create" else self.ann[mode])
        return F

    def string(self, modes) -> sp.csr_matrix:
        op = sp.identity(self.dim, dtype=np.complex128, format="csr")
        for p in sorted(int(m) for m in modes):
            op = op @ self.cre[p]
        return op

    def cluster(self, mode_rows, amplitudes) -> sp.csr_matrix:
        T = sp.csr_matrix((self.dim, self.dim), dtype=np.complex128)
        for row, a in zip(mode_rows, amplitudes):
            if a != 0.0:
                T = T + a * self.string(row[row >= 0])
        return T.tocsr()

    @staticmethod
    def _exp_apply(T, v, sign: float, order: int) -> np.ndarray:
        out = v.copy()
        term = v.copy()
        for k in range(1, order + 1):
            term = sign * (T @ term) / k
            out = out + term
        return out

    def transformed(self, T, v) -> np.ndarray:
        """``exp(-T) K exp(T) v``."""

        order = self.M // 2 + 2
        return self._exp_apply(T, self.K @ self._exp_apply(T, v, 1.0, order), -1.0, order)

    def vacuum(self) -> np.ndarray:
        v = np.zeros(self.dim, dtype=np.complex128)
        v[0] = 1.0
        return v

    @staticmethod
    def masks(mode_rows) -> np.ndarray:
        return np.array([sum(1 << int(p) for p in row[row >= 0]) for row in mode_rows], dtype=np.int64)

    def residual(self, coords, amplitudes) -> np.ndarray:
        T = self.cluster(coords.modes, amplitudes)
        state = self.transformed(T, self.vacuum())
        return -1j * state[self.masks(coords.modes)]

    def expectation(self, coords, amplitudes, factors) -> complex:
        """``<0| F_1 ... F_k exp(T) |0>`` with the complete exponential (no rank prefix)."""

        T = self.cluster(coords.modes, amplitudes)
        state = self._exp_apply(T, self.vacuum(), 1.0, self.M // 2 + 2)
        for components in reversed(factors):
            state = self.linear(components) @ state
        return complex(state[0])

    def sigma(self, static_coords, amplitudes, response_coords, x) -> np.ndarray:
        T = self.cluster(static_coords.modes, amplitudes)
        X = self.cluster(response_coords.modes, x)
        vac = self.vacuum()
        state = self.transformed(T, X @ vac) - X @ self.transformed(T, vac)
        return state[self.masks(response_coords.modes)]


# ----------------------------------------------------------------------------- physical ED


def physical_annihilators(num_orbitals: int) -> list[np.ndarray]:
    return [c.T.toarray() for c in jw_creators(num_orbitals)]


def physical_hamiltonian(model, c) -> np.ndarray:
    E = np.asarray(model.physical_E)
    dim = c[0].shape[0]
    H = np.zeros((dim, dim), dtype=np.complex128)
    for p in range(E.shape[0]):
        for q in range(E.shape[0]):
            if E[p, q] != 0.0:
                H += E[p, q] * c[p].conj().T @ c[q]
    for p, q, r, s, v in model.interaction.terms:
        H += v * c[p].conj().T @ c[q].conj().T @ c[s] @ c[r]
    return H


def lindblad_generators(model):
    """Return ``(rho_ss, L_even, L_odd, annihilators)`` in column-major vec."""

    N = model.n_spin_orbitals
    c = physical_annihilators(N)
    H = physical_hamiltonian(model, c)
    loss = np.asarray(model.physical_gamma_loss)
    gain = np.asarray(model.physical_gamma_gain)
    dim = H.shape[0]
    I = np.eye(dim, dtype=np.complex128)
    L = -1j * (np.kron(I, H) - np.kron(H.T, I))
    odd = np.zeros_like(L)
    for i in range(N):
        ci_dag = c[i].conj().T
        for j in range(N):
            ln = ci_dag @ c[j]
            L += loss[i, j] * (2.0 * np.kron(c[i].conj(), c[j]) - np.kron(I, ln) - np.kron(ln.T, I))
            gn = c[j] @ ci_dag
            L += gain[i, j] * (2.0 * np.kron(c[j].T, ci_dag) - np.kron(I, gn) - np.kron(gn.T, I))
            odd -= 4.0 * loss[i, j] * np.kron(c[i].conj(), c[j])
            odd -= 4.0 * gain[i, j] * np.kron(c[j].T, ci_dag)
    system = L.copy()
    rhs = np.zeros(dim * dim, dtype=np.complex128)
    system[0] = np.eye(dim).reshape(-1, order="F")
    rhs[0] = 1.0
    rho = np.linalg.solve(system, rhs).reshape((dim, dim), order="F")
    return rho, L, L + odd, c


def exact_retarded_green(model, omega, eta: float, orbital: int = 0) -> np.ndarray:
    """``G^R(ω) = Tr[c (z - i L_odd)^{-1} (c† ρ + ρ c†)]`` at ``z = ω + iη``."""

    rho, _, L_odd, c = lindblad_generators(model)
    dim = rho.shape[0]
    op = c[orbital]
    right = (op.conj().T @ rho + rho @ op.conj().T).reshape(-1, order="F")
    out = np.empty(len(omega), dtype=np.complex128)
    for k, w in enumerate(omega):
        A = (w + 1j * eta) * np.eye(dim * dim) - 1j * L_odd
        resp = np.linalg.solve(A, right).reshape((dim, dim), order="F")
        out[k] = np.trace(op @ resp)
    return out


def ph_neutral_spectrum(problem) -> np.ndarray:
    """Eigenvalues of the PH-mode ``K`` on the charge-neutral sector (``Q_up = Q_down = 0``)."""

    space = PHFockSpace(problem)
    n = problem.model.n_sites
    states = np.arange(space.dim)
    charge = np.zeros((space.dim, 2), dtype=np.int64)
    for p in range(space.M):
        spin, local = divmod(p, 2 * n)
        occ = (states >> p) & 1
        charge[:, spin] += np.where(local < n, occ, -occ)
    sel = np.flatnonzero((charge[:, 0] == 0) & (charge[:, 1] == 0))
    K = space.K.toarray()[np.ix_(sel, sel)]
    return np.linalg.eigvals(K)


def lindblad_superoperator_sparse(model):
    """``L_even`` of ``lindblad_generators`` (same column-major vec and the same terms), built
    from sparse Kronecker products, without the steady state; returns ``(L, dim)``."""

    import scipy.sparse as sparse

    N = model.n_spin_orbitals
    c = [sparse.csr_matrix(m) for m in physical_annihilators(N)]
    H = sparse.csr_matrix(physical_hamiltonian(model, [m.toarray() for m in c]))
    loss = np.asarray(model.physical_gamma_loss)
    gain = np.asarray(model.physical_gamma_gain)
    dim = H.shape[0]
    I = sparse.identity(dim, dtype=np.complex128, format="csr")
    L = -1j * (sparse.kron(I, H) - sparse.kron(H.T, I))
    for i in range(N):
        ci_dag = c[i].conj().T
        for j in range(N):
            if loss[i, j] != 0.0:
                ln = ci_dag @ c[j]
                L = L + loss[i, j] * (2.0 * sparse.kron(c[i].conj(), c[j]) - sparse.kron(I, ln) - sparse.kron(ln.T, I))
            if gain[i, j] != 0.0:
                gn = c[j] @ ci_dag
                L = L + gain[i, j] * (2.0 * sparse.kron(c[j].T, ci_dag) - sparse.kron(I, gn) - sparse.kron(gn.T, I))
    return L.tocsr(), dim


def lindblad_neutral_spectrum(model) -> np.ndarray:
    """Eigenvalues of ``L`` on operators ``|a><b|`` with equal up and down particle numbers."""

    L, dim = lindblad_superoperator_sparse(model)
    states = np.arange(dim)
    n_up = np.array([bin(int(s) & 0x5555555555555555).count("1") for s in states])
    n_dn = np.array([bin(int(s) & 0xAAAAAAAAAAAAAAAA).count("1") for s in states])
    # column-major vec index of |a><b| is a + dim*b
    a, b = np.meshgrid(states, states, indexing="ij")
    keep = (n_up[a] == n_up[b]) & (n_dn[a] == n_dn[b])
    sel = np.flatnonzero(keep.reshape(-1, order="F"))
    return np.linalg.eigvals(L[sel][:, sel].toarray())


def exact_covariance(model) -> np.ndarray:
    """``C[p, q] = Tr(c_q† c_p ρ)`` over physical spin orbitals."""

    rho, _, _, c = lindblad_generators(model)
    N = len(c)
    return np.array([[np.trace(c[q].conj().T @ c[p] @ rho) for q in range(N)] for p in range(N)])

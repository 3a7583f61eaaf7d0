"""Open fermionic impurity models: one-body matrices, Lindblad gain/loss, interactions.

Host-side mirror of the reference model types (``src/c3l/models.py:198-501``,
``src/c3l/interactions.py:15-139``, ``src/c3l/conventions.py:43-136``).  The
semantics that the coupled-cluster lowering depends on are kept identical:

* spin mode lifts every spatial matrix to site-major, spin-minor spin orbitals
  ``p = 2*site + spin`` (``models.py:519-528``);
* the interaction is stored as canonical antisymmetrised entries
  ``v[p,q,r,s]`` with ``p<q``, ``r<s`` meaning ``v a_p^† a_q^† a_s a_r``
  (``interactions.py:15-103``);
* a scalar ``U`` becomes ``U n_up n_down`` on every interaction site
  (``interactions.py:106-139``).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any

import numpy as np
from numpy.typing import ArrayLike, NDArray

ComplexMatrix = NDArray[np.complex128]


@dataclass(frozen=True)
class AbsRelTolerance:
    """``absolute + relative*|scale|`` threshold (``conventions.py:11-35``)."""

    absolute: float
    relative: float

    def threshold(self, scale: float) -> float:
        return float(self.absolute + self.relative * abs(float(scale)))


@dataclass(frozen=True)
class TolerancePolicy:
    """The reference default error budgets (``conventions.py:101-112``)."""

    input_abs_rel: AbsRelTolerance = AbsRelTolerance(1.0e-13, 1.0e-12)
    algebra_abs_rel: AbsRelTolerance = AbsRelTolerance(1.0e-13, 1.0e-12)
    eigensolver_abs_rel: AbsRelTolerance = AbsRelTolerance(1.0e-12, 1.0e-10)
    krylov_abs_rel: AbsRelTolerance = AbsRelTolerance(1.0e-12, 1.0e-10)
    nonlinear_abs_rel: AbsRelTolerance = AbsRelTolerance(1.0e-10, 1.0e-10)
    normal_mode_condition_budget_fraction: float = 0.1

    @classmethod
    def first(cls) -> "TolerancePolicy":
        return cls()


def _frozen(array: ArrayLike, dtype=np.complex128) -> NDArray:
    out = np.array(array, dtype=dtype, copy=True)
    out.flags.writeable = False
    return out


def _canonical_matrix(name: str, value: ArrayLike, tol: AbsRelTolerance, *, psd: bool) -> ComplexMatrix:
    """Validate a square Hermitian (optionally PSD) matrix, symmetrising roundoff.

    Mirrors ``models.py:56-118``: a Hermiticity defect above the input
    tolerance is an error, below it the matrix is symmetrised; tiny negative
    eigenvalues of a PSD input are clipped.
    """

    m = np.array(value, dtype=np.complex128, copy=True)
    if m.ndim != 2 or m.shape[0] != m.shape[1]:
        raise ValueError(f"{name} must be a square matrix")
    if m.size == 0:
        raise ValueError(f"{name} must be nonempty")
    if not np.all(np.isfinite(m)):
        raise ValueError(f"{name} must contain only finite values")
    threshold = tol.threshold(np.linalg.norm(m))
    defect = np.linalg.norm(m - m.conj().T)
    if defect > threshold:
        raise ValueError(f"{name} Hermiticity defect {defect:.3e} exceeds {threshold:.3e}")
    if defect > 0.0:
        m = 0.5 * (m + m.conj().T)
    if psd:
        w, v = np.linalg.eigh(m)
        psd_threshold = tol.threshold(np.linalg.norm(m, ord=2))
        if w[0] < -psd_threshold:
            raise ValueError(f"{name} minimum eigenvalue {w[0]:.3e} is below {-psd_threshold:.3e}")
        if w[0] < 0.0:
            m = (v * np.maximum(w, 0.0)) @ v.conj().T
            m = 0.5 * (m + m.conj().T)
    return _frozen(m)


TwoBodyTerm = tuple[int, int, int, int, complex]


@dataclass(frozen=True)
class SparseTwoBodyInteraction:
    """Canonical nonzero ``v[p,q,r,s]`` (``p<q``, ``r<s``) of ``1/4 v a†a†aa``.

    Same canonicalisation, sign handling and Hermiticity check as
    ``interactions.py:15-103``.
    """

    n_spin_orbitals: int
    terms: tuple[TwoBodyTerm, ...]

    def __post_init__(self) -> None:
        n = self.n_spin_orbitals
        if isinstance(n, bool) or not isinstance(n, int) or n < 1:
            raise ValueError("n_spin_orbitals must be a positive integer")
        combined: dict[tuple[int, int, int, int], complex] = {}
        for raw in self.terms:
            if not isinstance(raw, tuple) or len(raw) != 5:
                raise TypeError("each two-body term must be a five-element tuple")
            p, q, r, s, value = raw
            for index in (p, q, r, s):
                if isinstance(index, bool) or not isinstance(index, (int, np.integer)):
                    raise TypeError("two-body indices must be integers")
                if not 0 <= index < n:
                    raise ValueError("a two-body index lies outside the spin-orbital space")
            p, q, r, s = int(p), int(q), int(r), int(s)
            value = complex(value)
            if p == q or r == s:
                if value != 0.0:
                    raise ValueError("antisymmetric index pairs must contain distinct orbitals")
                continue
            if not (np.isfinite(value.real) and np.isfinite(value.imag)):
                raise ValueError("two-body coefficients must be finite")
            if p > q:
                p, q, value = q, p, -value
            if r > s:
                r, s, value = s, r, -value
            combined[(p, q, r, s)] = combined.get((p, q, r, s), 0.0j) + value
        canonical = tuple((*key, val) for key, val in sorted(combined.items()) if val != 0.0)
        lookup = {t[:4]: t[4] for t in canonical}
        for (p, q, r, s), val in lookup.items():
            partner = lookup.get((r, s, p, q))
            scale = max(abs(val), abs(partner) if partner is not None else 0.0, 1.0)
            if partner is None or abs(partner - val.conjugate()) > 1.0e-12 * scale:
                raise ValueError("two-body interaction must be Hermitian")
        object.__setattr__(self, "terms", canonical)


def density_density_interaction(
    n_sites: int, interaction_sites: tuple[int, ...], matrix: ArrayLike
) -> SparseTwoBodyInteraction:
    """``sum_i W_ii n_i↑ n_i↓ + sum_{i<j} W_ij n_i n_j`` (``interactions.py:106-139``)."""

    w = np.asarray(matrix, dtype=np.float64)
    terms: list[TwoBodyTerm] = []
    for a, si in enumerate(interaction_sites):
        up, down = 2 * si, 2 * si + 1
        if w[a, a] != 0.0:
            terms.append((up, down, up, down, complex(w[a, a])))
        for b in range(a + 1, len(interaction_sites)):
            sj = interaction_sites[b]
            if w[a, b] == 0.0:
                continue
            for s1 in (0, 1):
                for s2 in (0, 1):
                    p, q = 2 * si + s1, 2 * sj + s2
                    terms.append((p, q, p, q, complex(w[a, b])))
    return SparseTwoBodyInteraction(2 * n_sites, tuple(terms))


def spin_lift(matrix: ArrayLike) -> ComplexMatrix:
    """Site-major, spin-minor lift of a spatial matrix (``models.py:519-528``)."""

    a = np.asarray(matrix, dtype=np.complex128)
    out = np.zeros((2 * a.shape[0], 2 * a.shape[1]), dtype=np.complex128)
    out[0::2, 0::2] = a
    out[1::2, 1::2] = a
    return out


@dataclass(frozen=True)
class FermionicLindbladModel:
    """Number-conserving fermionic Lindblad model (spin symmetry).

    Mirrors ``models.py:198-417``.  ``symmetry="charge"`` is accepted by the
    reference; this build lowers spin-symmetric models only and rejects charge
    mode with a clear error.
    """

    E: ArrayLike
    gamma_loss: ArrayLike
    gamma_gain: ArrayLike
    U: float
    impurity_site: int = 0
    units: str = "Gamma"
    input_tolerance: AbsRelTolerance = field(default_factory=lambda: TolerancePolicy.first().input_abs_rel)
    interaction_sites: tuple[int, ...] | None = None
    density_interaction: ArrayLike | None = None
    two_body_interaction: SparseTwoBodyInteraction | None = None
    symmetry: str = "spin"

    def __post_init__(self) -> None:
        if self.symmetry not in ("spin", "charge"):
            raise ValueError("symmetry must be 'spin' or 'charge'")
        if self.symmetry == "charge":
            raise NotImplementedError("charge-symmetry models are not lowered by this build")
        tol = self.input_tolerance
        E = _canonical_matrix("E", self.E, tol, psd=False)
        loss = _canonical_matrix("gamma_loss", self.gamma_loss, tol, psd=True)
        gain = _canonical_matrix("gamma_gain", self.gamma_gain, tol, psd=True)
        if E.shape != loss.shape or E.shape != gain.shape:
            raise ValueError("E, gamma_loss, and gamma_gain must have identical shapes")
        raw_U = np.asarray(self.U)
        if raw_U.ndim != 0 or (np.iscomplexobj(raw_U) and np.imag(raw_U) != 0.0):
            raise ValueError("U must be a real scalar")
        U = float(np.real(raw_U))
        if not np.isfinite(U):
            raise ValueError("U must be finite")
        if isinstance(self.impurity_site, bool) or not isinstance(self.impurity_site, int):
            raise TypeError("impurity_site must be an integer")
        if not 0 <= self.impurity_site < E.shape[0]:
            raise ValueError("impurity_site is outside the spatial-site range")
        sites = (self.impurity_site,) if self.interaction_sites is None else tuple(self.interaction_sites)
        if not sites or len(set(sites)) != len(sites) or any(not 0 <= s < E.shape[0] for s in sites):
            raise ValueError("interaction_sites must be distinct sites inside the model")
        if sites[0] != self.impurity_site:
            raise ValueError("impurity_site must be the first interaction site")
        density = self.density_interaction
        two_body = self.two_body_interaction
        if two_body is not None:
            if not isinstance(two_body, SparseTwoBodyInteraction):
                raise TypeError("two_body_interaction must be a SparseTwoBodyInteraction")
            if U != 0.0 or density is not None:
                raise ValueError("two_body_interaction is mutually exclusive with U and density_interaction")
            if two_body.n_spin_orbitals != 2 * E.shape[0]:
                raise ValueError("two_body_interaction does not match the model spin-orbital space")
            allowed = {2 * s + spin for s in sites for spin in (0, 1)}
            if any(i not in allowed for t in two_body.terms for i in t[:4]):
                raise ValueError("two_body_interaction must be confined to interaction_sites")
        if density is not None:
            if U != 0.0:
                raise ValueError("U must be zero when density_interaction supplies the full matrix")
            d = np.asarray(density)
            if d.shape != (len(sites), len(sites)):
                raise ValueError("density_interaction must be indexed by interaction_sites")
            if np.iscomplexobj(d) and np.any(np.imag(d) != 0.0):
                raise ValueError("density_interaction must be real")
            d = np.asarray(np.real(d), dtype=np.float64)
            if not np.all(np.isfinite(d)):
                raise ValueError("density_interaction must be finite")
            if np.linalg.norm(d - d.T) > tol.threshold(max(np.linalg.norm(d), 1.0)):
                raise ValueError("density_interaction must be symmetric")
            density = _frozen(0.5 * (d + d.T), np.float64)
        object.__setattr__(self, "E", E)
        object.__setattr__(self, "gamma_loss", loss)
        object.__setattr__(self, "gamma_gain", gain)
        object.__setattr__(self, "U", U)
        object.__setattr__(self, "interaction_sites", sites)
        object.__setattr__(self, "density_interaction", density)

    @property
    def n_sites(self) -> int:
        return int(self.E.shape[0])

    @property
    def n_spin_orbitals(self) -> int:
        return 2 * self.n_sites

    @property
    def physical_E(self) -> ComplexMatrix:
        return spin_lift(self.E)

    @property
    def physical_gamma_loss(self) -> ComplexMatrix:
        return spin_lift(self.gamma_loss)

    @property
    def physical_gamma_gain(self) -> ComplexMatrix:
        return spin_lift(self.gamma_gain)

    def orbital_site_spin(self, orbital: int) -> tuple[int, int]:
        return divmod(int(orbital), 2)

    @property
    def interaction_matrix(self) -> NDArray[np.float64]:
        if self.density_interaction is not None:
            return np.asarray(self.density_interaction)
        return np.eye(len(self.interaction_sites)) * self.U

    @property
    def interaction(self) -> SparseTwoBodyInteraction:
        if self.two_body_interaction is not None:
            return self.two_body_interaction
        return density_density_interaction(self.n_sites, self.interaction_sites, self.interaction_matrix)

    @property
    def has_interactions(self) -> bool:
        return bool(self.interaction.terms)

    def to_dict(self) -> dict[str, Any]:
        """The reference's canonical model record (``models.py:428-452``)."""

        result: dict[str, Any] = {
            "model_type": "FermionicLindbladModel",
            "symmetry": self.symmetry,
            "E": _matrix_to_dict(self.E),
            "gamma_loss": _matrix_to_dict(self.gamma_loss),
            "gamma_gain": _matrix_to_dict(self.gamma_gain),
            "U": float(self.U),
            "impurity_site": self.impurity_site,
            "units": self.units,
            "input_tolerance": {"absolute": float(self.input_tolerance.absolute),
                                "relative": float(self.input_tolerance.relative)},
            "corrections": [],
        }
        if self.interaction_sites != (self.impurity_site,):
            result["interaction_sites"] = list(self.interaction_sites)
        if self.density_interaction is not None:
            result["density_interaction"] = np.asarray(self.density_interaction).tolist()
        if self.two_body_interaction is not None:
            result["two_body_interaction"] = {
                "n_spin_orbitals": self.two_body_interaction.n_spin_orbitals,
                "terms": [{"indices": [p, q, r, s], "coefficient": [float(v.real), float(v.imag)]}
                          for p, q, r, s, v in self.two_body_interaction.terms],
            }
        return result

    @classmethod
    def from_dict(cls, data) -> "FermionicLindbladModel":
        """Read the reference's model record (``models.py:454-501``).

        Matrices are ``{"real": ..., "imag": ...}`` records (plain nested lists are accepted
        too); ``corrections`` are derived from the matrices and ignored on input.
        """

        model_type = data.get("model_type", "FermionicLindbladModel")
        if model_type != "FermionicLindbladModel":
            raise ValueError(f"unsupported model_type {model_type!r}")
        tolerance_data = data.get("input_tolerance")
        tolerance = (TolerancePolicy.first().input_abs_rel if tolerance_data is None
                     else AbsRelTolerance(float(tolerance_data["absolute"]), float(tolerance_data["relative"])))
        sites = data.get("interaction_sites")
        two_body_data = data.get("two_body_interaction")
        two_body = None
        if two_body_data is not None:
            two_body = SparseTwoBodyInteraction(
                n_spin_orbitals=int(two_body_data["n_spin_orbitals"]),
                terms=tuple((*(int(i) for i in record["indices"]), complex(*record["coefficient"]))
                            for record in two_body_data["terms"]),
            )
        return cls(
            E=_matrix_from_dict(data["E"]),
            gamma_loss=_matrix_from_dict(data["gamma_loss"]),
            gamma_gain=_matrix_from_dict(data["gamma_gain"]),
            U=data["U"],
            impurity_site=int(data.get("impurity_site", 0)),
            units=str(data.get("units", "Gamma")),
            input_tolerance=tolerance,
            interaction_sites=None if sites is None else tuple(int(site) for site in sites),
            density_interaction=data.get("density_interaction"),
            two_body_interaction=two_body,
            symmetry=data.get("symmetry", "spin"),
        )


def _matrix_to_dict(matrix) -> dict[str, list]:
    m = np.asarray(matrix, dtype=np.complex128)
    return {"real": m.real.tolist(), "imag": m.imag.tolist()}


def _matrix_from_dict(data) -> np.ndarray:
    if isinstance(data, dict):
        return np.asarray(data["real"], dtype=np.float64) + 1j * np.asarray(data["imag"], dtype=np.float64)
    return np.asarray(data, dtype=np.complex128)

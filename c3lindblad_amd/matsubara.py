"""Fermionic Matsubara grids and the EOM impurity Green function on them.

Mirrors ``MatsubaraGrid``/``matsubara_grid`` (``src/c3l/matsubara/grid.py:16-210``) for
explicit grids and ``component_matsubara_green`` (``src/c3l/matsubara/solver.py:16-69``).
The retarded function is evaluated at the distinct upper-half-plane points ``i|omega_n|``
with one device Krylov space per source and channel
(``response.component_retarded_matrix_green_shifts``); negative frequencies take the
conjugate transpose, ``G(-i omega) = G(i omega)^dagger``.  DLR grids need ``pydlr``, which
this build does not use: ``dlr_matsubara_grid`` raises ``ImportError`` as the reference does
without it.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .response import component_retarded_matrix_green_shifts

_MATSUBARA_TOLERANCE = 1.0e-8


@dataclass(frozen=True, eq=False)
class MatsubaraGrid:
    """Points ``i omega_n``, ``omega_n = (2n+1) pi / beta``, in the caller's index order."""

    beta: float
    indices: np.ndarray
    kind: str = "explicit"
    _upper_odd: np.ndarray = field(init=False, repr=False)
    _upper_index: np.ndarray = field(init=False, repr=False)

    def __post_init__(self) -> None:
        beta = float(self.beta)
        if not np.isfinite(beta) or beta <= 0.0:
            raise ValueError("beta must be finite and positive")
        raw = np.asarray(self.indices)
        if raw.ndim != 1 or raw.size == 0 or not np.issubdtype(raw.dtype, np.integer):
            raise ValueError("indices must be a nonempty integer vector")
        indices = raw.astype(np.int64)
        if np.unique(indices).size != indices.size:
            raise ValueError("Matsubara indices must be unique")
        if self.kind != "explicit":
            raise ValueError("only explicit Matsubara grids are built (DLR grids need pydlr)")
        indices.setflags(write=False)
        odd, inverse = np.unique(np.abs(2 * indices + 1), return_inverse=True)
        object.__setattr__(self, "beta", beta)
        object.__setattr__(self, "indices", indices)
        object.__setattr__(self, "_upper_odd", odd)
        object.__setattr__(self, "_upper_index", inverse.astype(np.int64))

    def __len__(self) -> int:
        return int(self.indices.size)

    @property
    def omega_n(self) -> np.ndarray:
        return (2 * self.indices + 1) * np.pi / self.beta

    @property
    def iw(self) -> np.ndarray:
        return 1j * self.omega_n

    @property
    def upper_shifts(self) -> np.ndarray:
        """Distinct points ``i |omega_n|``, increasing."""

        return 1j * self._upper_odd * np.pi / self.beta

    def from_upper(self, values) -> np.ndarray:
        """Values at ``upper_shifts`` (scalar or square-matrix valued) mapped onto the grid."""

        upper = np.asarray(values, dtype=np.complex128)
        if upper.ndim not in (1, 3) or upper.shape[0] != self._upper_odd.size:
            raise ValueError("values must have shape (len(upper_shifts),) or (len(upper_shifts), m, m)")
        if upper.ndim == 3 and upper.shape[1] != upper.shape[2]:
            raise ValueError("matrix values must be square")
        out = upper[self._upper_index].copy()
        negative = self.indices < 0
        if upper.ndim == 1:
            out[negative] = out[negative].conj()
        else:
            out[negative] = out[negative].conj().swapaxes(1, 2)
        return out


def matsubara_grid(beta: float, *, indices=None, omega_n=None, n_iw: int | None = None) -> MatsubaraGrid:
    """Explicit fermionic grid from exactly one of ``indices``, ``omega_n`` or ``n_iw``
    (``grid.py:140-178``; ``n_iw`` gives ``n = -n_iw .. n_iw-1``)."""

    if sum(v is not None for v in (indices, omega_n, n_iw)) != 1:
        raise ValueError("give exactly one of indices, omega_n, or n_iw")
    b = float(beta)
    if not np.isfinite(b) or b <= 0.0:
        raise ValueError("beta must be finite and positive")
    if n_iw is not None:
        if isinstance(n_iw, bool) or not isinstance(n_iw, (int, np.integer)) or n_iw <= 0:
            raise ValueError("n_iw must be a positive integer")
        return MatsubaraGrid(b, np.arange(-int(n_iw), int(n_iw)))
    if indices is not None:
        return MatsubaraGrid(b, np.asarray(indices))
    frequency = np.asarray(omega_n)
    if np.iscomplexobj(frequency):
        if np.any(np.abs(frequency.real) > _MATSUBARA_TOLERANCE * np.maximum(np.abs(frequency), 1.0)):
            raise ValueError("complex Matsubara points must be purely imaginary")
        frequency = frequency.imag
    frequency = np.asarray(frequency, dtype=np.float64)
    if frequency.ndim != 1 or frequency.size == 0 or not np.all(np.isfinite(frequency)):
        raise ValueError("omega_n must be a finite nonempty vector")
    odd = frequency * b / np.pi
    nearest = np.rint(odd)
    if (np.any(np.abs(odd - nearest) > _MATSUBARA_TOLERANCE * np.maximum(np.abs(odd), 1.0))
            or np.any(nearest.astype(np.int64) % 2 == 0)):
        raise ValueError("omega_n are not fermionic Matsubara frequencies for this beta")
    return MatsubaraGrid(b, (nearest.astype(np.int64) - 1) // 2)


def dlr_matsubara_grid(beta: float, eps: float, *, lamb=None, wmax=None):
    raise ImportError("DLR grids require pydlr, which this build does not use")


@dataclass(frozen=True, eq=False)
class MatsubaraGreen:
    """Impurity ``G(i omega_n)`` (axes n, sink, source) (``solver.py:16-34``)."""

    grid: MatsubaraGrid
    sites: tuple
    green: np.ndarray
    upper: object

    @property
    def max_relative_residual(self) -> float:
        return float(max(np.max(self.upper.addition_relative_residuals), np.max(self.upper.removal_relative_residuals)))


def component_matsubara_green(
    operator,
    sources,
    grid: MatsubaraGrid,
    *,
    maximum_krylov_dimension: int = 320,
    residual_tolerance: float = 1.0e-10,
    minimum_krylov_dimension: int = 16,
    check_interval: int = 8,
) -> MatsubaraGreen:
    """EOM impurity Green function on a Matsubara grid (``solver.py:37-69``)."""

    if not isinstance(grid, MatsubaraGrid):
        raise TypeError("grid must be a MatsubaraGrid")
    upper = component_retarded_matrix_green_shifts(
        operator, sources, grid.upper_shifts, maximum_krylov_dimension=maximum_krylov_dimension,
        residual_tolerance=residual_tolerance, minimum_krylov_dimension=minimum_krylov_dimension,
        check_interval=check_interval,
    )
    return MatsubaraGreen(grid=grid, sites=upper.sites, green=grid.from_upper(upper.retarded), upper=upper)

"""Static observables of a converged CCSD state on the GPU.

Drop-in for ``factorized_ccsd_covariance`` (``src/c3l/cc/observables.py:164-200``)
and the rank-four expectations of ``examples/junction/observables.py:44-96``.

The reference forms the rank-0/2/4 prefix ``1 + T + T1∧T1/2`` of
``exp(T)|0>`` and dots the adjoint product with it
(``cc/observables.py:102-161``, ``ci/matrix_free.py:298-324``).  Here the
same number is read as the degree-0 part of ``exp(-T) O exp(T)|0>``: since
``<0|exp(-T) = <0|``, that part is ``<0|O exp(T)|0>``, and the term compiler
that builds the residual expands it into products of fully contracted ``T``
blocks.  Every observable of a request becomes one degree-0 block of one
``observable`` program, evaluated on the device in a single pass after the
amplitudes are loaded.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from time import perf_counter

import numpy as np

from . import _lib
from .algebra import TermCompiler
from .lowering import STATIC_KEYS, FactorizedModeProduct, physical_mode_operator
from .plan_arrays import LoweredPlan, lower_plan
from .static import ComponentStaticProblem


@dataclass
class ComponentObservableProblem:
    """Compiled expectation values ``<0| F_1 ... F_k exp(T) |0>`` of a list of products."""

    static: ComponentStaticProblem
    products: tuple[FactorizedModeProduct, ...]
    lowered: LoweredPlan
    compile_seconds: float
    _device: object = field(default=None, repr=False)

    @classmethod
    def compile(cls, static: ComponentStaticProblem, products) -> "ComponentObservableProblem":
        from .algebra import emit_plan

        products = tuple(products)
        if not products:
            raise ValueError("at least one observable product is required")
        for record in products:
            if not isinstance(record, FactorizedModeProduct):
                raise TypeError("observables are FactorizedModeProduct records")
            if len(record.factors) > 4:
                raise ValueError("the rank-0/2/4 prefix supports products of at most four linear operators")
        started = perf_counter()
        oracle = static.oracle
        compiler = TermCompiler(oracle.model.n_sites, products)
        scalars = [compiler.product_terms(c, factors, (), 0) for c, factors in compiler.products]
        plan = emit_plan({}, oracle.model.n_sites, compiler.table.array(), None, scalars)
        lowered = lower_plan(plan, {"T": (oracle.coordinates, STATIC_KEYS)}, {})
        return cls(static, products, lowered, perf_counter() - started)

    @property
    def device(self):
        if self._device is None:
            from .device import DevicePlan

            self._device = DevicePlan(self.lowered)
        return self._device

    def evaluate(self, amplitudes) -> np.ndarray:
        """All expectations at the amplitudes (NumPy complex or a device tensor)."""

        torch = _lib.require_device()
        plan = self.device
        if isinstance(amplitudes, np.ndarray) or not hasattr(amplitudes, "is_cuda"):
            t = np.ascontiguousarray(amplitudes, dtype=np.complex128)
            if t.shape != (self.static.oracle.num_amplitudes,) or not np.all(np.isfinite(t)):
                raise ValueError("amplitudes must be finite and match the cluster basis")
            td = torch.from_numpy(t).to(_lib.DEVICE)
        else:
            td = amplitudes
        out = torch.zeros(len(self.products), dtype=torch.complex128, device=_lib.DEVICE)
        plan.load("T", td)
        plan.apply("observable", out)
        return out.cpu().numpy()


def _operators(reference):
    n = reference.model.n_spin_orbitals
    spin = [reference.model.orbital_site_spin(o) for o in range(n)]
    ann = tuple(physical_mode_operator(reference, s, p, "c") for s, p in spin)
    cre = tuple(physical_mode_operator(reference, s, p, "c_dagger") for s, p in spin)
    return ann, cre


def covariance_problem(static: ComponentStaticProblem) -> ComponentObservableProblem:
    """The compiled ``<c_q† c_p>`` request for every pair of spin orbitals (cached on ``static``)."""

    cached = getattr(static, "_covariance_problem", None)
    if cached is None:
        ann, cre = _operators(static.oracle.reference)
        n = len(ann)
        products = [FactorizedModeProduct("CCSD_covariance", 1.0, (cre[q], ann[p])) for p in range(n) for q in range(n)]
        cached = ComponentObservableProblem.compile(static, products)
        static._covariance_problem = cached
    return cached


def factorized_ccsd_covariance(static: ComponentStaticProblem, amplitudes) -> np.ndarray:
    """``C[p, q] = <c_q† c_p>`` in physical orbital order (``cc/observables.py:164-200``).

    Values are not symmetrised or clipped, so numerical defects stay visible.
    """

    values = covariance_problem(static).evaluate(amplitudes)
    n = static.oracle.model.n_spin_orbitals
    return values.reshape(n, n)


def two_body_expectations(static: ComponentStaticProblem, amplitudes, quartets) -> np.ndarray:
    """``<c_p† c_q c_r† c_s>`` for every ``(p, q, r, s)`` of ``quartets`` (spin-orbital indices).

    Same contraction as ``four(p, q, r, s)`` of ``examples/junction/observables.py:82-83``.
    """

    ann, cre = _operators(static.oracle.reference)
    n = len(ann)
    products = []
    for quartet in quartets:
        p, q, r, s = (int(i) for i in quartet)
        if not all(0 <= i < n for i in (p, q, r, s)):
            raise ValueError("spin-orbital index outside the model")
        products.append(FactorizedModeProduct("two_body", 1.0, (cre[p], ann[q], cre[r], ann[s])))
    return ComponentObservableProblem.compile(static, products).evaluate(amplitudes)


def contact_currents(model, gamma, impurity, reservoirs: dict) -> dict:
    """Reservoir-to-impurity currents and dissipative reservoir fluxes from the covariance.

    ``I_r = 2 Im sum_{p in impurity, q in r} h[p, q] gamma[q, p]`` and
    ``F_r = 2 Re(tr gain[r, r] - sum_{q in r, k} (loss + gain)[q, k] gamma[k, q])``
    over spin orbitals (``examples/junction/model.py:150-175``, which fixes the
    impurity to the first eight spin orbitals and two reservoirs; here both are
    arguments).  ``gamma[p, q] = <c_q† c_p>``.
    """

    gamma = np.asarray(gamma, dtype=np.complex128)
    N = model.n_spin_orbitals
    if gamma.shape != (N, N) or not np.all(np.isfinite(gamma)):
        raise ValueError("gamma must use the complete finite spin-orbital space")
    h = model.physical_E
    loss = model.physical_gamma_loss
    gain = model.physical_gamma_gain
    imp = np.asarray(impurity, dtype=np.int64)
    out = {}
    for label, orbitals in reservoirs.items():
        q = np.asarray(orbitals, dtype=np.int64)
        out[f"I_{label}"] = float(2.0 * np.sum(h[np.ix_(imp, q)] * gamma[np.ix_(q, imp)].T).imag)
        flux = np.trace(gain[np.ix_(q, q)]) - np.sum((loss + gain)[q, :] * gamma[:, q].T)
        out[f"F_{label}"] = float(2.0 * flux.real)
    return out

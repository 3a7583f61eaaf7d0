"""EOM-CCSD (S1+S3) response space, context, sigma operator, sources and sink.

Drop-in for ``eom_ccsd_problem`` (``src/c3l/cc/eom.py:93-135``),
``ComponentEOMProblem.compile`` (``cc/component_tensor.py:2012-2056``),
``component_eom_context`` (``:2234-2251``), ``component_eom_sigma``
(``:2288-2302``), ``component_eom_sources`` (``:2780-2858``) and
``ComponentSigmaOperator`` (``cc/component_response.py:28-58``).

``σ(X) = Λ∘X + P_13[exp(-T) K exp(T) X|0> - X exp(-T) K exp(T)|0>]`` in one
charge block of response spin 0 (no ``-i``: ``K = iL``).  The commutator
keeps exactly the terms in which the input is contracted at least once, so
the compiler emits only those.  A context uploads the plan, loads the
converged amplitudes once (their atoms stay resident) and every matvec
rebuilds only the input atoms.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from time import perf_counter

import numpy as np

from . import _lib
from .algebra import TermCompiler, emit_plan, factor_linear_terms, term_key
from .lowering import STATIC_KEYS, CoordinateSet, FactorizedCCProblem, physical_mode_operator, response_coordinates, response_keys
from .plan_arrays import LoweredPlan, lower_plan
from .static import ComponentStaticProblem


@dataclass
class EOMCCSDProblem:
    cc: FactorizedCCProblem
    response_spin: int = 0

    @property
    def dimension(self) -> int:
        n = self.cc.model.n_sites
        return 2 * (n + n * (n * (n - 1) // 2) + n**3)


def eom_ccsd_problem(cc: FactorizedCCProblem, *, response_spin: int = 0) -> EOMCCSDProblem:
    if not isinstance(cc, FactorizedCCProblem):
        raise TypeError("cc must be a factorized CC problem")
    if response_spin != 0:
        raise ValueError("this build compiles response_spin=0 (the reference default)")
    return EOMCCSDProblem(cc, response_spin)


@dataclass
class ComponentEOMProblem:
    oracle: EOMCCSDProblem
    static: ComponentStaticProblem
    charge: int
    coordinates: CoordinateSet
    keys: tuple
    lowered: LoweredPlan
    diagonal: np.ndarray

    @classmethod
    def compile(cls, problem: EOMCCSDProblem, static: ComponentStaticProblem, charge: int) -> "ComponentEOMProblem":
        if problem.cc is not static.oracle:
            raise ValueError("static and EOM component problems do not share a CC problem")
        if charge not in (-1, 1):
            raise ValueError("response charge must be +1 or -1")
        n = problem.cc.model.n_sites
        coords = response_coordinates(n, charge)
        keys = tuple(b.key for b in coords.blocks)
        compiler = TermCompiler(n, problem.cc.products)
        terms = {k: v for k, v in compiler.response_terms(keys, 3).items() if term_key(k) in keys}
        factored, creations = factor_linear_terms(terms, "X")
        plan = emit_plan({"sigma": (factored, keys)}, n, compiler.table.array(), creations)
        lowered = lower_plan(plan, {"T": (static.oracle.coordinates, STATIC_KEYS), "X": (coords, keys)}, {"sigma": coords})
        diagonal = problem.cc.rapidity_sums(coords)
        return cls(problem, static, charge, coords, keys, lowered, diagonal)

    @property
    def dimension(self) -> int:
        return self.coordinates.dimension

    def device_plan(self, slot: int = 0):
        """Uploaded sigma plan number ``slot``; independent slots can run concurrently."""

        plans = self.__dict__.setdefault("_plans", {})
        if slot not in plans:
            from .device import DevicePlan

            plans[slot] = DevicePlan(self.lowered)
        return plans[slot]


@dataclass
class ComponentEOMContext:
    problem: ComponentEOMProblem
    amplitudes: np.ndarray
    _devices: dict = field(default_factory=dict, repr=False)

    @property
    def device(self) -> "EOMDevice":
        return self.device_for(0)

    def device_for(self, slot: int) -> "EOMDevice":
        """Device state on plan slot ``slot`` (amplitude atoms loaded on the current stream)."""

        if slot not in self._devices:
            self._devices[slot] = EOMDevice(self, slot)
        return self._devices[slot]


def component_eom_context(problem: ComponentEOMProblem, amplitudes) -> ComponentEOMContext:
    t = np.asarray(amplitudes, dtype=np.complex128)
    if t.shape != (problem.static.oracle.num_amplitudes,) or not np.all(np.isfinite(t)):
        raise ValueError("amplitudes must be finite and match the CCSD basis")
    context = ComponentEOMContext(problem, t)
    context.device  # make the amplitude atoms resident on plan slot 0
    return context


class EOMDevice:
    """Device state of one context: the problem's plan with this context's amplitude atoms."""

    def __init__(self, context: ComponentEOMContext, slot: int = 0):
        torch = _lib.require_device()
        self.torch = torch
        self.plan = context.problem.device_plan(slot)
        self.t = torch.from_numpy(np.ascontiguousarray(context.amplitudes)).to(_lib.DEVICE)
        self.token = self.plan.load("T", self.t)
        self.diag = torch.from_numpy(np.ascontiguousarray(context.problem.diagonal)).to(_lib.DEVICE)
        self.dim = context.problem.dimension

    def sigma(self, x, out) -> None:
        """``out = σ(x)`` for contiguous complex device vectors of the block dimension."""

        if self.plan.loaded.get(0) != self.token:
            self.token = self.plan.load("T", self.t)
        self.plan.load("X", x)
        self.plan.apply("sigma", out, self.diag, x)


@dataclass
class ComponentEOMAction:
    retained: np.ndarray
    wall_seconds: float


def component_eom_sigma(context: ComponentEOMContext, vector) -> ComponentEOMAction:
    started = perf_counter()
    x = np.asarray(vector, dtype=np.complex128)
    if x.shape != (context.problem.dimension,) or not np.all(np.isfinite(x)):
        raise ValueError("EOM vector has the wrong shape or is nonfinite")
    dev = context.device
    xd = dev.torch.from_numpy(np.ascontiguousarray(x)).to(_lib.DEVICE)
    out = dev.torch.empty_like(xd)
    dev.sigma(xd, out)
    return ComponentEOMAction(out.cpu().numpy(), perf_counter() - started)


class ComponentSigmaOperator:
    """Instrumented matrix-free sigma (``component_response.py:43-58``).

    ``matvec`` keeps the NumPy contract (complex128 in/out); ``matvec_device``
    applies it to device tensors without host traffic.
    """

    def __init__(self, context: ComponentEOMContext, *, slot: int = 0):
        self.context = context
        self.slot = slot
        self.dimension = context.problem.dimension
        self.diagonal = context.problem.diagonal.copy()
        self.action_calls = 0
        self.action_wall_seconds = 0.0

    @property
    def device(self) -> "EOMDevice":
        return self.context.device_for(self.slot)

    def on_slot(self, slot: int) -> "ComponentSigmaOperator":
        """Same operator bound to another plan slot (for concurrent resolvents)."""

        return ComponentSigmaOperator(self.context, slot=slot)

    def matvec(self, vector) -> np.ndarray:
        started = perf_counter()
        result = component_eom_sigma(self.context, vector).retained
        self.action_calls += 1
        self.action_wall_seconds += perf_counter() - started
        return np.asarray(result, dtype=np.complex128)

    def matvec_device(self, x, out) -> None:
        self.device.sigma(x, out)
        self.action_calls += 1


@dataclass
class ComponentEOMSources:
    addition: np.ndarray
    removal: np.ndarray
    sink: np.ndarray
    charge: int
    site: int


def _block_values(coords: CoordinateSet, key, values: np.ndarray, out: np.ndarray) -> None:
    block = coords.block(key)
    if block is not None and block.positions.size:
        out[block.positions] += values[tuple(block.indices.T)]


def _source(problem: ComponentEOMProblem, amplitudes: np.ndarray, components) -> np.ndarray:
    """``pack(F_creator + ι_F T)`` (``component_tensor.py:2788-2797``), host FP64."""

    n = problem.static.oracle.model.n_sites
    coords = problem.coordinates
    out = np.zeros(coords.dimension, dtype=np.complex128)
    create = np.zeros((4, n), dtype=np.complex128)
    annihilate = np.zeros((4, n), dtype=np.complex128)
    for kind, mode, c in components:
        g, q = divmod(mode, n)
        (create if kind == "create" else annihilate)[g, q] += c
    for g in range(4):
        key = tuple(1 if k == g else 0 for k in range(4))
        _block_values(coords, key, create[g], out)
    # interior of the amplitudes with the annihilator part, group by group
    static_coords = problem.static.oracle.coordinates
    from .plan_arrays import padded_scatter

    for key in STATIC_KEYS:
        src, sign = padded_scatter(static_coords, key)
        deg = sum(key)
        padded = (np.where(src >= 0, amplitudes[np.maximum(src, 0)], 0.0) * sign).reshape((n,) * deg)
        slots = [g for g, c in enumerate(key) for _ in range(c)]
        for g in range(4):
            if key[g] == 0 or not np.any(annihilate[g]):
                continue
            s = slots.index(g)
            value = np.tensordot(padded, annihilate[g], axes=([s], [0])) * (-1.0 if s % 2 else 1.0)
            rest = list(key)
            rest[g] -= 1
            _block_values(coords, tuple(rest), value, out)
    return out


def _sink(problem: ComponentEOMProblem, components) -> np.ndarray:
    n = problem.static.oracle.model.n_sites
    out = np.zeros(problem.dimension, dtype=np.complex128)
    for kind, mode, c in components:
        if kind != "annihilate":
            continue
        g, q = divmod(mode, n)
        key = tuple(1 if k == g else 0 for k in range(4))
        block = problem.coordinates.block(key)
        if block is not None:
            hit = block.indices[:, 0] == q
            out[block.positions[hit]] = c
    return out


def component_eom_sources(context: ComponentEOMContext, *, site: int | None = None) -> ComponentEOMSources:
    """Addition/removal sources and the trace sink (``component_tensor.py:2816-2858``)."""

    cc = context.problem.static.oracle
    target = cc.model.impurity_site if site is None else site
    if isinstance(target, (bool, np.bool_)) or not isinstance(target, (int, np.integer)) or not 0 <= int(target) < cc.model.n_sites:
        raise ValueError("site is outside the auxiliary model")
    target = int(target)
    spin = context.problem.oracle.response_spin
    ref = cc.reference
    t = context.amplitudes
    if context.problem.charge == 1:
        addition = _source(context.problem, t, physical_mode_operator(ref, target, spin, "c_dagger").components)
        removal = -1j * _source(context.problem, t, physical_mode_operator(ref, target, spin, "tilde_c").components)
        sink_op = physical_mode_operator(ref, target, spin, "c")
    else:
        addition = -1j * _source(context.problem, t, physical_mode_operator(ref, target, spin, "tilde_c_dagger").components)
        removal = _source(context.problem, t, physical_mode_operator(ref, target, spin, "c").components)
        sink_op = physical_mode_operator(ref, target, spin, "c_dagger")
    return ComponentEOMSources(addition, removal, _sink(context.problem, sink_op.components), context.problem.charge, target)

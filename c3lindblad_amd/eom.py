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
    space: str = "S13"  # "S135": S1 + S3 + the whole rank-five shell

    @property
    def dimension(self) -> int:
        """Response coordinates of both charge blocks."""

        n = self.cc.model.n_sites
        return sum(response_coordinates(n, q, self.space).dimension for q in (1, -1))

    @property
    def max_rank(self) -> int:
        return 5 if self.space == "S135" else 3


def eom_ccsd_problem(cc: FactorizedCCProblem, *, response_spin: int = 0, space: str = "S13") -> EOMCCSDProblem:
    """``space="S135"`` adds the rank-five shell (the reference's full S1 + S3 + S5 response,
    ``cc/s135_full.py:101-361``) to the S13 space of ``cc/eom.py:93-135``."""

    if not isinstance(cc, FactorizedCCProblem):
        raise TypeError("cc must be a factorized CC problem")
    if response_spin != 0:
        raise ValueError("this build compiles response_spin=0 (the reference default)")
    if space not in ("S13", "S135"):
        raise ValueError("response space must be 'S13' or 'S135'")
    return EOMCCSDProblem(cc, response_spin, space)


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
        coords = response_coordinates(n, charge, problem.space)
        keys = tuple(b.key for b in coords.blocks)
        compiler = TermCompiler(n, problem.cc.products)
        terms = {k: v for k, v in compiler.response_terms(keys, problem.max_rank).items() if term_key(k) in keys}
        factored, creations = factor_linear_terms(terms, "X")
        plan = emit_plan({"sigma": (factored, keys)}, n, compiler.table.array(), creations)
        lowered = lower_plan(plan, {"T": (static.oracle.coordinates, STATIC_KEYS), "X": (coords, keys)}, {"sigma": coords})
        diagonal = problem.cc.rapidity_sums(coords)
        return cls(problem, static, charge, coords, keys, lowered, diagonal)

    @property
    def dimension(self) -> int:
        return self.coordinates.dimension

    def source_plan(self, site: int) -> LoweredPlan:
        """Plan whose ``source_addition``/``source_removal`` programs evaluate the two sources of ``site``.

        A source is ``pack(F_creator + ι_F T)`` (``component_tensor.py:2788-2797``): the
        degree-1 and degree-3 part of ``exp(-T) F exp(T)|0>`` for the channel operator ``F``
        of the charge (``component_tensor.py:2816-2858``), so the term compiler expands it
        like one factor of a residual product.
        """

        plans = self.__dict__.setdefault("_source_plans", {})
        if site not in plans:
            cc = self.oracle.cc
            ops = channel_operators(cc.reference, site, self.oracle.response_spin, self.charge)
            compiler = TermCompiler(cc.model.n_sites, ())
            actions = {}
            for name in ("addition", "removal"):
                op, coefficient = ops[name]
                terms = compiler.apply_factor({(): complex(coefficient)}, compiler._lin(op), self.oracle.max_rank)
                actions["source_" + name] = ({k: v for k, v in terms.items() if term_key(k) in self.keys}, self.keys)
            plan = emit_plan(actions, cc.model.n_sites, compiler.table.array())
            plans[site] = lower_plan(
                plan,
                {"T": (cc.coordinates, STATIC_KEYS)},
                {"source_addition": self.coordinates, "source_removal": self.coordinates},
            )
        return plans[site]

    def source_device(self, site: int):
        """Uploaded source plan of ``site`` (shared by every context of this problem)."""

        plans = self.__dict__.setdefault("_source_devices", {})
        if site not in plans:
            from .device import DevicePlan

            plans[site] = DevicePlan(self.source_plan(site))
        return plans[site]

    def sink(self, site: int) -> np.ndarray:
        """Rank-one annihilator coefficients of the channel's sink operator (cached per site)."""

        sinks = self.__dict__.setdefault("_sinks", {})
        if site not in sinks:
            cc = self.oracle.cc
            op = channel_operators(cc.reference, site, self.oracle.response_spin, self.charge)["sink"][0]
            sinks[site] = _sink(self, op.components)
        return sinks[site]

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


def _check_device_amplitudes(device_amplitudes, shape) -> None:
    """The plan reads the amplitudes as contiguous complex128: anything else is rejected on the host."""

    if (getattr(device_amplitudes, "is_cuda", None) is None or tuple(device_amplitudes.shape) != tuple(shape)
            or str(device_amplitudes.dtype) != "torch.complex128" or not device_amplitudes.is_contiguous()):
        raise ValueError("device_amplitudes must be a contiguous complex128 GPU vector matching the CCSD basis")
    if not device_amplitudes.is_cuda:
        raise ValueError("device_amplitudes must be a contiguous complex128 GPU vector matching the CCSD basis")


def component_eom_context(problem: ComponentEOMProblem, amplitudes, *, device_amplitudes=None) -> ComponentEOMContext:
    """``device_amplitudes`` (a GPU copy of the same amplitudes, e.g. from the solver) skips the upload."""

    t = np.asarray(amplitudes, dtype=np.complex128)
    if t.shape != (problem.static.oracle.num_amplitudes,) or not np.all(np.isfinite(t)):
        raise ValueError("amplitudes must be finite and match the CCSD basis")
    context = ComponentEOMContext(problem, t)
    if device_amplitudes is not None:
        _check_device_amplitudes(device_amplitudes, t.shape)
        context.__dict__["_device_t"] = device_amplitudes
    context.device  # make the amplitude atoms resident on plan slot 0
    return context


class EOMDevice:
    """Device state of one context: the problem's plan with this context's amplitude atoms."""

    def __init__(self, context: ComponentEOMContext, slot: int = 0):
        torch = _lib.require_device()
        self.torch = torch
        self.plan = context.problem.device_plan(slot)
        shared = context.__dict__.get("_device_t")
        if shared is None:
            shared = torch.from_numpy(np.ascontiguousarray(context.amplitudes)).to(_lib.DEVICE)
            context.__dict__["_device_t"] = shared
        self.t = shared
        self.token = self.plan.load("T", self.t)
        diag = context.problem.__dict__.get("_device_diagonal")
        if diag is None:
            diag = torch.from_numpy(np.ascontiguousarray(context.problem.diagonal)).to(_lib.DEVICE)
            context.problem.__dict__["_device_diagonal"] = diag
        self.diag = diag
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


def channel_operators(reference, site: int, spin: int, charge: int) -> dict:
    """Source operators (with their scalar factors) and sink of one charge channel.

    Charge +1: addition ``c†``, removal ``-i c̃``, sink ``c``; charge -1: addition
    ``-i c̃†``, removal ``c``, sink ``c†`` (``component_tensor.py:2829-2851``).
    """

    if charge == 1:
        return {
            "addition": (physical_mode_operator(reference, site, spin, "c_dagger"), 1.0),
            "removal": (physical_mode_operator(reference, site, spin, "tilde_c"), -1j),
            "sink": (physical_mode_operator(reference, site, spin, "c"), 1.0),
        }
    return {
        "addition": (physical_mode_operator(reference, site, spin, "tilde_c_dagger"), -1j),
        "removal": (physical_mode_operator(reference, site, spin, "c"), 1.0),
        "sink": (physical_mode_operator(reference, site, spin, "c_dagger"), 1.0),
    }


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
    """Addition/removal sources and the trace sink (``component_tensor.py:2816-2858``), sources on the GPU."""

    cc = context.problem.static.oracle
    target = cc.model.impurity_site if site is None else site
    if isinstance(target, (bool, np.bool_)) or not isinstance(target, (int, np.integer)) or not 0 <= int(target) < cc.model.n_sites:
        raise ValueError("site is outside the auxiliary model")
    target = int(target)
    problem = context.problem
    plan = problem.source_device(target)
    torch = _lib.require_device()
    dev = context.device
    plan.load("T", dev.t)
    add = torch.zeros(problem.dimension, dtype=torch.complex128, device=_lib.DEVICE)
    rem = torch.zeros_like(add)
    plan.apply("source_addition", add)
    plan.apply("source_removal", rem)
    return ComponentEOMSources(add.cpu().numpy(), rem.cpu().numpy(), problem.sink(target), problem.charge, target)

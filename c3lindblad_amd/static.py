"""CCSD steady-state residual and Jacobian-vector products on the GPU.

Drop-in for ``ComponentStaticProblem.compile``, ``component_static_residual``,
``component_static_linearization``, ``component_static_linearized_jvp`` and
``component_static_jvp`` (``src/c3l/cc/component_tensor.py:1276-1298,
1371-1549,1552-1657,1822-1868,1871-1930``).

``R(T) = -i (Λ∘T + sum_products [exp(-T) prod F exp(T)|0>]_{retained})`` with
``Λ_μ = sum_{p in μ} λ_p``; the rank-0 part is the ``reference`` scalar.

Compile (host, once per problem): the term compiler expands all products,
differentiates them for the tangent, and the plan is uploaded.  Each action
is then: scatter the input into padded base blocks, contract the atoms
level by level, evaluate every output block at its canonical coordinates.
The linearisation keeps the ``T`` atoms resident in the plan workspace, so a
JVP only rebuilds the tangent atoms.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from time import perf_counter

import numpy as np

from . import _lib
from .algebra import TermCompiler, emit_plan, factor_linear_terms, tangent_terms, term_key
from .lowering import STATIC_KEYS, FactorizedCCProblem
from .plan_arrays import LoweredPlan, lower_plan


def _t_degree(term_key) -> int:
    """Amplitude atoms of a term: its degree in the amplitudes."""

    return sum(1 for atom in term_key if getattr(atom, "base", None) == "T")


@dataclass
class ComponentStaticAction:
    retained: np.ndarray
    reference: complex
    wall_seconds: float


@dataclass
class ComponentStaticProblem:
    """Compiled static CCSD problem (host plan + lazily created device state)."""

    oracle: FactorizedCCProblem
    lowered: LoweredPlan
    diagonal: np.ndarray  # -i Λ_μ
    compile_seconds: float
    residual_degree: int = 4  # highest power of the amplitudes in any residual term
    _device: "StaticDevice | None" = field(default=None, repr=False)

    @classmethod
    def compile(cls, problem: FactorizedCCProblem, *, fuse_products: bool = False) -> "ComponentStaticProblem":
        started = perf_counter()
        compiler = TermCompiler(problem.model.n_sites, problem.products)
        residual = compiler.static_terms(4)
        jvp = {k: -1j * v for k, v in tangent_terms(residual).items() if term_key(k) in STATIC_KEYS}
        jvp_factored, creations = factor_linear_terms(jvp, "D")
        # R(0) + J(0) t / 6 without the diagonal term, the end point s = 0 of Simpson's rule: the
        # terms with no amplitude atom, and a sixth of those with one, factored as constant
        # creation tensors (products of the generator's atoms alone) times one amplitude atom
        low = {k: -1j * v for k, v in residual.items() if _t_degree(k) == 0 and term_key(k) in STATIC_KEYS}
        linear = {k: -1j * v / 6.0 for k, v in residual.items() if _t_degree(k) == 1 and term_key(k) in STATIC_KEYS}
        low_factored, low_creations = factor_linear_terms(linear, "T", first_index=len(creations))
        overlap = set(low) & set(low_factored)
        if overlap:
            raise AssertionError("constant and factored linear residual terms collide")
        low.update(low_factored)
        plan = emit_plan(
            {
                "residual": ({k: -1j * v for k, v in residual.items()}, STATIC_KEYS),
                "reference": ({k: -1j * v for k, v in residual.items()}, ((0, 0, 0, 0),)),
                "jvp": (jvp_factored, STATIC_KEYS),
                "residual_low": (low, STATIC_KEYS),
            },
            problem.model.n_sites,
            compiler.table.array(),
            creations + low_creations,
        )
        coords = problem.coordinates
        lowered = lower_plan(
            plan,
            {"T": (coords, STATIC_KEYS), "D": (coords, STATIC_KEYS)},
            {"residual": coords, "reference": coords, "jvp": coords, "residual_low": coords},
        )
        diagonal = -1j * problem.rapidity_sums()
        degree = max((_t_degree(key) for key in residual), default=0)
        return cls(problem, lowered, diagonal, perf_counter() - started, max(int(degree), 1))

    @property
    def coordinates(self):
        return self.oracle.coordinates

    @property
    def device(self) -> "StaticDevice":
        if self._device is None:
            self._device = StaticDevice(self)
        return self._device


class StaticDevice:
    """GPU buffers and actions of one compiled static problem."""

    def __init__(self, problem: ComponentStaticProblem):
        from .device import DevicePlan, DeviceTilde

        torch = _lib.require_device()
        self.torch = torch
        self.problem = problem
        self.plan = DevicePlan(problem.lowered)
        self.tilde = DeviceTilde(problem.oracle.tilde)
        dim = problem.oracle.num_amplitudes
        self.dim = dim
        self.real_dim = problem.oracle.num_variables
        dev = torch.device(_lib.DEVICE)
        self.diag = torch.from_numpy(np.ascontiguousarray(problem.diagonal)).to(dev)
        self.t = torch.zeros(dim, dtype=torch.complex128, device=dev)
        self.dt = torch.zeros(dim, dtype=torch.complex128, device=dev)
        self.out = torch.zeros(dim, dtype=torch.complex128, device=dev)
        self.ref = torch.zeros(1, dtype=torch.complex128, device=dev)
        self.t_token = 0
        # residual evaluation: "program" (the residual program), "quadrature" (FP64 JVPs at
        # Gauss-Legendre points) or "simpson" (FP64 JVPs at s = 1/2 and 1 plus the low-degree
        # program); the set-up tuner picks the fastest for the problem
        self.residual_path = "program"
        self._r0 = None  # R(0), the constant part, for the quadrature path
        self._ts = None
        self._qout = None
        self._diag6 = None  # diag / 6, the diagonal's share of the Simpson end point s = 0

    # complex-vector actions ------------------------------------------------
    def load_amplitudes(self, t) -> None:
        if t.data_ptr() != self.t.data_ptr():
            self.t.copy_(t)
        self.t_token = self.plan.load("T", self.t)

    def residual_at(self, t, out=None):
        """Residual at the amplitudes ``t`` (complex, on the device) into ``out``; the amplitude
        atoms of ``t`` are loaded afterwards.  The quadrature and Simpson paths load the atoms at
        their own nodes and end with ``t``, so only the program path loads ``t`` first."""

        if t.data_ptr() != self.t.data_ptr():
            self.t.copy_(t)
        if self.residual_path == "program":
            self.t_token = self.plan.load("T", self.t)
        return self.residual_complex(out)

    def residual_complex(self, out=None):
        """Residual of the currently loaded amplitudes into ``out`` (complex)."""

        out = self.out if out is None else out
        if self.residual_path == "quadrature":
            self._residual_by_quadrature(out)
        elif self.residual_path == "simpson":
            self._residual_by_simpson(out)
        else:
            self.plan.apply("residual", out, self.diag, self.t)
        self.plan.apply("reference", self.ref)
        return out

    def quadrature_nodes(self) -> tuple[np.ndarray, np.ndarray]:
        """Gauss-Legendre nodes and weights on [0, 1] that integrate ``s -> J(s t) t`` exactly: a
        residual of degree d makes it a polynomial of degree d - 1, and k nodes are exact up to
        degree 2k - 1."""

        k = max(1, (self.problem.residual_degree + 1) // 2)
        x, w = np.polynomial.legendre.leggauss(k)
        return 0.5 * (x + 1.0), 0.5 * w

    def _residual_by_quadrature(self, out) -> None:
        """``R(t) = R(0) + int_0^1 J(s t) t ds``: the residual's multi-factor terms (the costly
        gathers of the residual program) become FP64 JVPs, whose amplitude-only factors are the
        creation tensors materialised at each node.  Exact for the polynomial residual (to
        rounding); the amplitude atoms of ``t`` are loaded again at the end."""

        torch = self.torch
        if self._ts is None:
            self._ts = torch.empty_like(self.t)
            self._qout = torch.empty_like(self.t)
        if self._r0 is None:
            self._ts.zero_()
            self.plan.load("T", self._ts)
            self._r0 = torch.empty_like(self.t)
            self.plan.apply("residual", self._r0, self.diag, self._ts)
        bits = self.plan.precision("jvp")
        if bits != 64:
            self.plan.set_precision("D", "jvp", 64)
        try:
            out.copy_(self._r0)
            for s, w in zip(*self.quadrature_nodes()):
                torch.mul(self.t, float(s), out=self._ts)
                self.plan.load("T", self._ts)
                self.plan.load("D", self.t)
                self.plan.apply("jvp", self._qout, self.diag, self.t)
                out.add_(self._qout, alpha=float(w))
        finally:
            if bits != 64:
                self.plan.set_precision("D", "jvp", bits)
            self.t_token = self.plan.load("T", self.t)

    def _residual_by_simpson(self, out) -> None:
        """Simpson's rule on ``R(t) = R(0) + int_0^1 f(s) ds`` with ``f(s) = J(s t) t``, a cubic in
        ``s`` for a residual of degree at most four, so the rule is exact (to rounding):
        ``R(t) = R(0) + f(0)/6 + 4 f(1/2)/6 + f(1)/6``.  ``f(1/2)`` and ``f(1)`` are FP64 JVPs with
        the amplitude atoms loaded at ``t/2`` and then at ``t``; ``R(0) + f(0)/6`` is the
        ``residual_low`` program (terms of degree zero, and a sixth of the linear terms as constant
        creation tensors times one amplitude atom) with a sixth of the diagonal, launched while
        ``t`` is loaded.  Two amplitude loads with their creation tensors instead of the three of
        two-node Gauss-Legendre, and the amplitudes end loaded at ``t`` as every residual path
        leaves them."""

        if self.problem.residual_degree > 4:
            raise ValueError(f"Simpson's rule is exact up to residual degree 4, not {self.problem.residual_degree}")
        torch = self.torch
        if self._ts is None:
            self._ts = torch.empty_like(self.t)
            self._qout = torch.empty_like(self.t)
        if self._diag6 is None:
            self._diag6 = self.diag / 6.0
        bits = self.plan.precision("jvp")
        if bits != 64:
            self.plan.set_precision("D", "jvp", 64)
        try:
            torch.mul(self.t, 0.5, out=self._ts)
            self.plan.load("T", self._ts)
            self.plan.load("D", self.t)
            self.plan.apply("jvp", self._qout, self.diag, self.t)  # f(1/2)
            self.t_token = self.plan.load("T", self.t)
            self.plan.apply("jvp", out, self.diag, self.t)  # f(1)
            out.add_(self._qout, alpha=4.0)
            self.plan.apply("residual_low", self._qout, self._diag6, self.t)  # R(0) + f(0)/6
            out.div_(6.0).add_(self._qout)
        finally:
            if bits != 64:
                self.plan.set_precision("D", "jvp", bits)

    def jvp_complex(self, dt, out=None):
        out = self.out if out is None else out
        if dt.data_ptr() != self.dt.data_ptr():
            self.dt.copy_(dt)
        self.plan.load("D", self.dt)
        self.plan.apply("jvp", out, self.diag, self.dt)
        return out

    # real tilde-coordinate actions -----------------------------------------
    def residual_real(self, x, r) -> None:
        self.tilde.unpack(x, self.t)
        self.residual_at(self.t, self.out)
        self.tilde.pack(self.out, r)

    def jvp_real(self, dx, jdx) -> None:
        self.tilde.unpack(dx, self.dt)
        self.plan.load("D", self.dt)
        self.plan.apply("jvp", self.out, self.diag, self.dt)
        self.tilde.pack(self.out, jdx)


def _to_device(array, torch):
    return torch.from_numpy(np.ascontiguousarray(array, dtype=np.complex128)).to(_lib.DEVICE)


def _validate(problem: ComponentStaticProblem, values, name: str) -> np.ndarray:
    a = np.asarray(values, dtype=np.complex128)
    if a.shape != (problem.oracle.num_amplitudes,) or not np.all(np.isfinite(a)):
        raise ValueError(f"{name} must be finite and match the CCSD basis")
    return a


def component_static_residual(problem: ComponentStaticProblem, amplitudes) -> ComponentStaticAction:
    started = perf_counter()
    t = _validate(problem, amplitudes, "amplitudes")
    dev = problem.device
    torch = dev.torch
    out = dev.residual_at(_to_device(t, torch))
    retained = out.cpu().numpy().copy()
    reference = complex(dev.ref.cpu().numpy()[0])
    return ComponentStaticAction(retained, reference, perf_counter() - started)


@dataclass
class ComponentStaticLinearization:
    """Primal state of one Newton step; its ``T`` atoms stay on the device."""

    problem: ComponentStaticProblem
    amplitudes: np.ndarray
    residual: ComponentStaticAction
    token: int


def component_static_linearization(problem: ComponentStaticProblem, amplitudes) -> ComponentStaticLinearization:
    action = component_static_residual(problem, amplitudes)
    return ComponentStaticLinearization(problem, _validate(problem, amplitudes, "amplitudes"), action, problem.device.t_token)


def _ensure_loaded(context: ComponentStaticLinearization) -> None:
    dev = context.problem.device
    if dev.t_token != context.token:
        dev.load_amplitudes(_to_device(context.amplitudes, dev.torch))
        context.token = dev.t_token


def component_static_linearized_jvp(context: ComponentStaticLinearization, tangent) -> ComponentStaticAction:
    started = perf_counter()
    dt = _validate(context.problem, tangent, "tangent")
    _ensure_loaded(context)
    dev = context.problem.device
    out = dev.jvp_complex(_to_device(dt, dev.torch))
    return ComponentStaticAction(out.cpu().numpy().copy(), 0.0j, perf_counter() - started)


def component_static_jvp(problem: ComponentStaticProblem, amplitudes, direction) -> ComponentStaticAction:
    context = component_static_linearization(problem, amplitudes)
    return component_static_linearized_jvp(context, direction)

"""Shift-invariant Krylov resolvent and retarded Green functions on the GPU.

Drop-in for ``component_retarded_green`` (``src/c3l/cc/component_response.py:180-245``)
and ``arnoldi_resolvent`` (``src/c3l/green/krylov.py:284-417``).

One Arnoldi basis per source serves every frequency, because the Krylov space
of ``(z - σ)`` does not depend on ``z``.  The native driver (``c3l_arnoldi``)
builds the basis with classical Gram--Schmidt with delayed reorthogonalisation
(one reduction and one update pass per step) and, for every shift at once, runs
GMRES on it: a Givens QR of ``(z I - H)`` advanced one column per
step and the sink projection ``G(z) = sink·x_m(z)`` updated incrementally.
Every shift's true relative residual ``||b - (z - σ) x_m(z)|| / ||b||`` is
exact at every step, so the basis stops growing exactly when the worst shift
meets the tolerance.  ``c3l_arnoldi_batch`` runs the addition and removal
bases of a charge block in lockstep on twin plans, one launch per kernel for
both.  The reference solves the same systems with one restarted
GMRES per frequency (``component_response.py:72-128``); both converge to the
same ``G^R`` within the stated tolerance.
"""

from __future__ import annotations

import ctypes
import resource
from dataclasses import dataclass
from time import perf_counter

import numpy as np
import scipy.linalg

from . import _lib
from .device import Reducer
from .eom import ComponentEOMSources, ComponentSigmaOperator


@dataclass
class ArnoldiResolvent:
    """Krylov projection ``σ V_m = V_m H_m + h v_{m+1} e_m^T`` with the basis kept on the GPU.

    The reference's ``ArnoldiResolvent`` (``green/krylov.py:28-254``): the same fields and
    methods (``coefficients``, ``residual_norms``, ``relative_residuals``, ``solutions``,
    ``time_coefficients``, ``time_residual_norms``, ``relative_time_residuals``, ``transfer``,
    ``to_dict``).  ``basis`` is a torch ``[m+1, P]`` complex128 device tensor whose rows are
    the basis vectors (the reference stores them as columns).  The reduced solves use the
    Schur form of ``H_m`` on the host, as the reference does; projections of sinks on the
    basis and reconstructed solutions run on the device.  ``shifts``/``green``/
    ``shift_relative_residuals`` are the native multi-shift results for the shifts and sink
    the basis was built with (GMRES or Galerkin, following ``criterion``).
    ``action_wall_seconds`` holds the construction time spread evenly over the actions (the
    native driver does not time actions one by one); ``check_dimensions`` holds the final
    dimension only.
    """

    basis: object  # torch [m+1, P] complex128 (rows are basis vectors)
    hessenberg: np.ndarray  # (m+1, m)
    source_norm: float
    breakdown: bool
    shifts: np.ndarray
    green: np.ndarray
    shift_relative_residuals: np.ndarray
    converged: bool
    construction_wall_seconds: float
    action_calls: int
    residual_tolerance: float = 0.0
    criterion: str = "gmres"
    peak_rss_kib: int = 0

    @property
    def dimension(self) -> int:
        return int(self.hessenberg.shape[1])

    @property
    def vector_dimension(self) -> int:
        return int(self.basis.shape[1])

    @property
    def terminal_coupling(self) -> complex:
        m = self.dimension
        return complex(self.hessenberg[m, m - 1])

    @property
    def action_wall_seconds(self) -> np.ndarray:
        m = max(self.dimension, 1)
        return np.full(self.dimension, float(self.construction_wall_seconds) / m)

    @property
    def check_dimensions(self) -> tuple:
        checks = self.__dict__.get("_checks")
        return checks[0] if checks is not None else (self.dimension,)

    @property
    def check_maximum_relative_residuals(self) -> np.ndarray:
        checks = self.__dict__.get("_checks")
        return checks[1] if checks is not None else np.asarray([float(np.max(self.shift_relative_residuals))])

    @property
    def orthogonality_defect(self) -> float:
        """``||V_m^H V_m - I||_F / m`` (``krylov.py:395-402``), one Gram product on the device."""

        cached = self.__dict__.get("_orthogonality")
        if cached is None:
            m = self.dimension
            V = self.basis[:m].contiguous()
            gram = (V.conj() @ V.transpose(0, 1).contiguous()).cpu().numpy()
            cached = float(np.linalg.norm(gram - np.eye(m), ord="fro") / max(float(m), 1.0))
            self.__dict__["_orthogonality"] = cached
        return cached

    def _dimension(self, dimension) -> int:
        m = self.dimension if dimension is None else dimension
        if isinstance(m, bool) or not isinstance(m, (int, np.integer)):
            raise TypeError("dimension must be an integer")
        if not 1 <= m <= self.dimension:
            raise ValueError("dimension must lie in [1, krylov_dimension]")
        return int(m)

    def _schur(self, m: int):
        key = ("_schur", m)
        cached = self.__dict__.get(key)
        if cached is None:
            T, Q = scipy.linalg.schur(self.hessenberg[:m, :m], output="complex", check_finite=False)
            cached = (T, Q)
            self.__dict__[key] = cached
        return cached

    def coefficients(self, shifts, *, dimension=None) -> np.ndarray:
        """Galerkin coordinates ``y(z) = (z - H_m)^{-1} beta e_1`` (``krylov.py:66-93``), shifts x m."""

        z = np.asarray(shifts, dtype=np.complex128)
        if z.ndim != 1 or z.size == 0 or not np.all(np.isfinite(z)):
            raise ValueError("shifts must be a finite nonempty vector")
        m = self._dimension(dimension)
        T, Q = self._schur(m)
        reduced_rhs = Q.conj().T[:, 0] * self.source_norm
        identity = np.eye(m, dtype=np.complex128)
        reduced = np.empty((z.size, m), dtype=np.complex128)
        for p, shift in enumerate(z):
            reduced[p] = Q @ scipy.linalg.solve_triangular(shift * identity - T, reduced_rhs, lower=False,
                                                           check_finite=False)
        return reduced

    def residual_norms(self, shifts, *, dimension=None) -> np.ndarray:
        m = self._dimension(dimension)
        reduced = self.coefficients(shifts, dimension=m)
        return np.asarray(np.abs(self.hessenberg[m, m - 1] * reduced[:, -1]), dtype=np.float64)

    def relative_residuals(self, shifts, *, dimension=None) -> np.ndarray:
        return np.asarray(self.residual_norms(shifts, dimension=dimension)
                          / max(float(self.source_norm), np.finfo(np.float64).tiny), dtype=np.float64)

    def _device_combination(self, reduced: np.ndarray, m: int):
        torch = _lib.require_device()
        y = torch.from_numpy(np.ascontiguousarray(reduced)).to(_lib.DEVICE)
        return y @ self.basis[:m].contiguous()

    def solutions(self, shifts, *, dimension=None) -> np.ndarray:
        """``V_m y(z)`` for every shift (shifts x P), formed on the device."""

        m = self._dimension(dimension)
        reduced = self.coefficients(shifts, dimension=m)
        return self._device_combination(reduced, m).cpu().numpy()

    def projected_sink(self, sink, m: int) -> np.ndarray:
        """``sink · V_m`` (no conjugation), one reduction over the basis on the device."""

        torch = _lib.require_device()
        row = np.asarray(sink, dtype=np.complex128)
        if row.shape != (self.vector_dimension,) or not np.all(np.isfinite(row)):
            raise ValueError("sink must match the full vector dimension")
        red = Reducer(self.vector_dimension, m, complex_=True)
        sink_d = torch.from_numpy(np.ascontiguousarray(row)).to(_lib.DEVICE)
        sv = torch.empty(m, dtype=torch.complex128, device=_lib.DEVICE)
        red.dot(self.basis, m, sink_d, sv, conjugate=False)
        return sv.cpu().numpy()

    def transfer(self, shifts, sink, *, dimension=None) -> np.ndarray:
        """``sink · V_m y(z)`` (``krylov.py:214-227``)."""

        m = self._dimension(dimension)
        projected = self.projected_sink(sink, m)
        return np.asarray(self.coefficients(shifts, dimension=m) @ projected, dtype=np.complex128)

    def time_coefficients(self, time, *, dimension=None) -> np.ndarray:
        """``exp(-i H_m t) beta e_1`` on a time grid (``krylov.py:136-172``)."""

        t = np.asarray(time, dtype=np.float64)
        if t.ndim != 1 or t.size == 0 or not np.all(np.isfinite(t)):
            raise ValueError("time must be a finite nonempty vector")
        m = self._dimension(dimension)
        # exp(-i t H_m) = Q exp(-i t T) Q^H with the Schur form H_m = Q T Q^H
        T, Q = self._schur(m)
        r = Q.conj().T[:, 0] * self.source_norm
        out = np.empty((t.size, m), dtype=np.complex128)
        for i, value in enumerate(t):
            out[i] = Q @ (scipy.linalg.expm(-1j * float(value) * T) @ r)
        return out

    def time_residual_norms(self, time, *, dimension=None) -> np.ndarray:
        """``|h_{m+1,m} e_m^T exp(-i H_m t) beta e_1|`` (``krylov.py:174-196``)."""

        m = self._dimension(dimension)
        reduced = self.time_coefficients(time, dimension=m)
        return np.asarray(np.abs(self.hessenberg[m, m - 1] * reduced[:, -1]), dtype=np.float64)

    def relative_time_residuals(self, time, *, dimension=None) -> np.ndarray:
        return np.asarray(self.time_residual_norms(time, dimension=dimension)
                          / max(float(self.source_norm), np.finfo(np.float64).tiny), dtype=np.float64)

    def project(self, sink, shifts):
        """``sink·V_m (z - H_m)^{-1} β e_1`` for other sinks/shifts (batched triangular solves on the
        device) and the Galerkin residual estimate ``|h_{m+1,m} e_m^T y| / β`` of every shift."""

        m = self.dimension
        T, Q = self._schur(m)
        r = Q.conj().T[:, 0] * self.source_norm
        w = self.projected_sink(sink, m) @ Q
        z = np.ascontiguousarray(np.asarray(shifts, dtype=np.complex128))
        green, last = shifted_green(T, r, w, Q[m - 1, :], z)
        residual = np.abs(self.hessenberg[m, m - 1] * last) / max(self.source_norm, np.finfo(float).tiny)
        return green, residual

    def to_dict(self) -> dict:
        """The reference's record (``krylov.py:229-254``)."""

        return {
            "krylov_dimension": self.dimension,
            "full_vector_dimension": self.vector_dimension,
            "source_norm": float(self.source_norm),
            "terminal_coupling_absolute": float(abs(self.terminal_coupling)),
            "orthogonality_defect": float(self.orthogonality_defect),
            "action_wall_seconds": [float(v) for v in self.action_wall_seconds],
            "action_wall_seconds_total": float(np.sum(self.action_wall_seconds)),
            "construction_wall_seconds": float(self.construction_wall_seconds),
            "peak_rss_kib": int(self.peak_rss_kib),
            "breakdown": bool(self.breakdown),
            "check_dimensions": list(self.check_dimensions),
            "check_maximum_relative_residuals": [float(v) for v in self.check_maximum_relative_residuals],
            "maximum_relative_residual": float(self.check_maximum_relative_residuals[-1]),
            "residual_tolerance": float(self.residual_tolerance),
            "converged": bool(self.converged),
            "method": "delayed_reorthogonalisation_CGS_shift_invariant_resolvent_" + self.criterion,
        }


def shifted_green(T, r, w, q, z):
    """Batched ``w·(z-T)^{-1} r`` and ``q·(z-T)^{-1} r`` on the GPU (upper-triangular T)."""

    torch = _lib.require_device()
    m = T.shape[0]

    def dev(a):
        return torch.from_numpy(np.ascontiguousarray(a, dtype=np.complex128)).to(_lib.DEVICE)

    # every operand is held by a name until the results are copied back
    Tc, rd, wd, qd, zd = dev(np.asfortranarray(T).ravel(order="F")), dev(r), dev(w), dev(q), dev(z)
    green = torch.empty(z.size, dtype=torch.complex128, device=_lib.DEVICE)
    last = torch.empty(z.size, dtype=torch.complex128, device=_lib.DEVICE)
    _lib.call(
        "c3l_shifted_triangular_green",
        _lib.dptr(Tc), m, _lib.dptr(rd), _lib.dptr(wd), _lib.dptr(qd), _lib.dptr(zd), z.size,
        _lib.dptr(green), _lib.dptr(last), _lib.stream_ptr(torch),
    )
    out = green.cpu().numpy(), last.cpu().numpy()
    del Tc, rd, wd, qd, zd
    return out


def _krylov_dimension_checks(maximum_dimension, minimum_dimension, check_interval, breakdown_tolerance, n):
    if isinstance(maximum_dimension, bool) or not isinstance(maximum_dimension, (int, np.integer)) \
            or not 0 < maximum_dimension <= n:
        raise ValueError("maximum_dimension must lie in [1, operator.dimension]")
    if isinstance(minimum_dimension, bool) or not isinstance(minimum_dimension, (int, np.integer)) \
            or not 1 <= minimum_dimension <= maximum_dimension:
        raise ValueError("minimum_dimension must lie in [1, maximum_dimension]")
    if isinstance(check_interval, bool) or not isinstance(check_interval, (int, np.integer)) or check_interval <= 0:
        raise ValueError("check_interval must be a positive integer")
    if not (np.isfinite(breakdown_tolerance) and breakdown_tolerance > 0.0):
        raise ValueError("breakdown_tolerance must be finite and positive")


def arnoldi_resolvent(
    operator: ComponentSigmaOperator,
    rhs,
    shifts,
    *,
    maximum_dimension: int,
    residual_tolerance: float,
    minimum_dimension: int = 16,
    check_interval: int = 8,
    breakdown_tolerance: float = 1.0e-14,
    sink=None,
    criterion: str = "galerkin",
) -> ArnoldiResolvent:
    """Build one shift-invariant Krylov space for ``(zI - σ)^{-1} b`` with the native driver
    (``arnoldi_resolvent``, ``green/krylov.py:284-417``).

    The basis stops growing at the first dimension ``m >= minimum_dimension`` at which every
    shift's Galerkin residual ``|h_{m+1,m} e_m^T y_m(z)| / ||b||`` is at most
    ``residual_tolerance`` (the reference's rule, tested here after every column rather than
    every ``check_interval`` columns, so the dimension is the smallest one that passes), at a
    breakdown, or at ``maximum_dimension``.  ``criterion="gmres"`` stops on the GMRES residual
    instead (the per-frequency solves of ``component_retarded_green``).  With ``sink`` the
    result's ``green`` holds the native GMRES value ``sink·x(z)`` for every shift.
    """

    torch = _lib.require_device()
    if not hasattr(operator, "dimension"):
        raise TypeError("operator must expose dimension and matvec")
    n = int(operator.dimension)
    b = np.asarray(rhs, dtype=np.complex128)
    z = np.ascontiguousarray(np.asarray(shifts, dtype=np.complex128))
    if b.shape != (n,) or not np.all(np.isfinite(b)):
        raise ValueError("rhs must be a finite vector matching the operator")
    if z.ndim != 1 or z.size == 0 or not np.all(np.isfinite(z)):
        raise ValueError("shifts must be a finite nonempty vector")
    _krylov_dimension_checks(maximum_dimension, minimum_dimension, check_interval, breakdown_tolerance, n)
    if not (np.isfinite(residual_tolerance) and residual_tolerance > 0.0):
        raise ValueError("residual_tolerance must be finite and positive")
    if criterion not in ("galerkin", "gmres"):
        raise ValueError("criterion must be 'galerkin' or 'gmres'")
    sink_h = np.zeros(n, dtype=np.complex128) if sink is None else np.asarray(sink, dtype=np.complex128)
    if sink_h.shape != (n,) or not np.all(np.isfinite(sink_h)):
        raise ValueError("sink must be a finite vector matching the operator")
    beta = float(np.linalg.norm(b))
    if beta == 0.0:
        raise ValueError("rhs must be nonzero")
    started = perf_counter()
    m_max = int(maximum_dimension)
    dev = _lib.DEVICE
    c128, f64 = torch.complex128, torch.float64
    ns = z.size
    # columns padded to c3l_krylov_basis_ld(n) entries (each starts on a cache line), written before read
    V = torch.empty(m_max + 1, _lib.krylov_basis_ld(n), dtype=c128, device=dev)
    V[0, :n].copy_(torch.from_numpy(b / beta))
    H = torch.zeros(m_max, m_max + 1, dtype=c128, device=dev)  # row j = column j of H
    zs = torch.from_numpy(z).to(dev)
    sink_d = torch.from_numpy(np.ascontiguousarray(sink_h)).to(dev)
    rot_c = torch.empty(ns * m_max, dtype=f64, device=dev)  # per-shift state, written before read
    rot_s = torch.empty(ns * m_max, dtype=c128, device=dev)
    pi = torch.empty(ns * m_max, dtype=c128, device=dev)
    gamma = torch.zeros(ns, dtype=c128, device=dev)
    green = torch.zeros(ns, dtype=c128, device=dev)
    resid = torch.zeros(ns, dtype=f64, device=dev)
    scratch_size, coef_size = _lib.krylov_workspace(True, n, m_max, ns)
    scratch = torch.empty(scratch_size, dtype=f64, device=dev)
    coef = torch.zeros(coef_size, dtype=f64, device=dev)
    status = torch.zeros(4, dtype=f64, device=dev)
    m_out, broke = ctypes.c_int32(), ctypes.c_int32()
    worst = ctypes.c_double()
    state = operator.device
    if state.plan.loaded.get(0) != state.token:
        state.token = state.plan.load("T", state.t)
    _lib.call(
        "c3l_arnoldi", state.plan.handle, _lib.dptr(state.diag), _lib.dptr(V), _lib.dptr(H), n, m_max, _lib.dptr(zs), ns,
        _lib.dptr(sink_d), beta, _lib.dptr(rot_c), _lib.dptr(rot_s), _lib.dptr(pi), _lib.dptr(gamma), _lib.dptr(green),
        _lib.dptr(resid), _lib.dptr(scratch), _lib.dptr(coef), _lib.dptr(status), float(residual_tolerance),
        int(minimum_dimension), float(breakdown_tolerance), 1 if criterion == "galerkin" else 0, ctypes.byref(m_out),
        ctypes.byref(broke), ctypes.byref(worst), _lib.stream_ptr(torch),
    )
    m = m_out.value
    operator.action_calls += m
    residuals = resid.cpu().numpy()
    return ArnoldiResolvent(
        basis=V[: m + 1, :n],
        hessenberg=H[:m, : m + 1].cpu().numpy().T.copy(),
        source_norm=beta,
        breakdown=bool(broke.value),
        shifts=z,
        green=green.cpu().numpy(),
        shift_relative_residuals=residuals,
        converged=bool(broke.value or np.max(residuals) <= residual_tolerance),
        construction_wall_seconds=perf_counter() - started,
        action_calls=m,
        residual_tolerance=float(residual_tolerance),
        criterion=criterion,
        peak_rss_kib=int(resource.getrusage(resource.RUSAGE_SELF).ru_maxrss),
    )


class _ArnoldiRun:
    """Device buffers and native run (``c3l_arnoldi_begin`` .. ``c3l_arnoldi_end``) of one
    ``arnoldi_resolvent_batch``, started on the current stream."""

    def __init__(self, operators, rhs, shifts, *, maximum_dimension, residual_tolerance, sinks,
                 minimum_dimension=16, breakdown_tolerance=1.0e-14):
        torch = _lib.require_device()
        count = len(operators)
        if not 1 <= count <= 4 or len(rhs) != count or len(sinks) != count:
            raise ValueError("between one and four operators, right-hand sides and sinks are required")
        contexts = {id(op.context) for op in operators}
        slots = [op.slot for op in operators]
        if len(contexts) != 1 or len(set(slots)) != count:
            raise ValueError("operators must share one context on distinct plan slots")
        n = int(operators[0].dimension)
        z = np.ascontiguousarray(np.asarray(shifts, dtype=np.complex128))
        if z.ndim != 1 or z.size == 0 or not np.all(np.isfinite(z)):
            raise ValueError("shifts must be a finite nonempty vector")
        if (isinstance(maximum_dimension, bool) or not isinstance(maximum_dimension, int)
                or not 0 < maximum_dimension <= n):
            raise ValueError("maximum_dimension must lie in [1, operator.dimension]")
        if not (np.isfinite(residual_tolerance) and residual_tolerance > 0.0):
            raise ValueError("residual_tolerance must be finite and positive")
        b = np.stack([np.asarray(v, dtype=np.complex128) for v in rhs])
        snk = np.stack([np.asarray(v, dtype=np.complex128) for v in sinks])
        if (b.shape != (count, n) or snk.shape != (count, n)
                or not (np.all(np.isfinite(b)) and np.all(np.isfinite(snk)))):
            raise ValueError("right-hand sides and sinks must be finite vectors matching the operator")
        beta = np.array([float(np.linalg.norm(v)) for v in b])  # rounded as in arnoldi_resolvent
        if np.any(beta == 0.0):
            raise ValueError("right-hand sides must be nonzero")
        self.started = perf_counter()
        self.operators, self.count, self.z, self.beta = operators, count, z, beta
        self.tolerance = float(residual_tolerance)
        m_max = maximum_dimension
        dev = _lib.DEVICE
        c128, f64 = torch.complex128, torch.float64
        ns = z.size
        # columns padded to c3l_krylov_basis_ld(n) entries (each starts on a cache line), written before read
        self.n = n
        self.V = torch.empty(count, m_max + 1, _lib.krylov_basis_ld(n), dtype=c128, device=dev)
        self.V[:, 0, :n].copy_(torch.from_numpy(b / beta[:, None]))
        self.H = torch.zeros(count, m_max, m_max + 1, dtype=c128, device=dev)
        zs = torch.from_numpy(z).to(dev)
        sink_d = torch.from_numpy(np.ascontiguousarray(snk)).to(dev)
        rot_c = torch.empty(count, ns * m_max, dtype=f64, device=dev)
        rot_s = torch.empty(count, ns * m_max, dtype=c128, device=dev)
        pi = torch.empty(count, ns * m_max, dtype=c128, device=dev)
        gamma = torch.zeros(count, ns, dtype=c128, device=dev)
        self.green = torch.zeros(count, ns, dtype=c128, device=dev)
        self.resid = torch.zeros(count, ns, dtype=f64, device=dev)
        scratch_size, coef_size = _lib.krylov_workspace(True, n, m_max, ns)
        scratch = torch.empty(count, scratch_size, dtype=f64, device=dev)
        coef = torch.zeros(count, coef_size, dtype=f64, device=dev)
        status = torch.zeros(count, 4, dtype=f64, device=dev)
        self._keep = (zs, sink_d, rot_c, rot_s, pi, gamma, scratch, coef, status)  # alive until the run ends
        states = [op.device for op in operators]
        for state in states:
            if state.plan.loaded.get(0) != state.token:
                state.token = state.plan.load("T", state.t)
        plans = (ctypes.c_void_p * count)(*[state.plan.handle for state in states])
        beta_h = (ctypes.c_double * count)(*beta.tolist())
        self.handle = ctypes.c_void_p()
        _lib.call(
            "c3l_arnoldi_begin", plans, count, _lib.dptr(states[0].diag), _lib.dptr(self.V), _lib.dptr(self.H), n,
            m_max, _lib.dptr(zs), ns, _lib.dptr(sink_d), beta_h, _lib.dptr(rot_c), _lib.dptr(rot_s), _lib.dptr(pi),
            _lib.dptr(gamma), _lib.dptr(self.green), _lib.dptr(self.resid), _lib.dptr(scratch), _lib.dptr(coef),
            _lib.dptr(status), self.tolerance, int(minimum_dimension), float(breakdown_tolerance),
            _lib.stream_ptr(torch), ctypes.byref(self.handle),
        )

    def finish(self) -> list[ArnoldiResolvent]:
        """End the run (waits for its stream) and return one projection per right-hand side."""

        count = self.count
        m_out = (ctypes.c_int32 * count)()
        broke = (ctypes.c_int32 * count)()
        worst = (ctypes.c_double * count)()
        handle, self.handle = self.handle, None
        _lib.call("c3l_arnoldi_end", handle, m_out, broke, worst)
        green_h = self.green.cpu().numpy()
        resid_h = self.resid.cpu().numpy()
        wall = perf_counter() - self.started
        out = []
        for i, op in enumerate(self.operators):
            m = int(m_out[i])
            op.action_calls += m
            out.append(ArnoldiResolvent(
                basis=self.V[i, : m + 1, : self.n],
                hessenberg=self.H[i, :m, : m + 1].cpu().numpy().T.copy(),
                source_norm=float(self.beta[i]),
                breakdown=bool(broke[i]),
                shifts=self.z,
                green=green_h[i].copy(),
                shift_relative_residuals=resid_h[i].copy(),
                converged=bool(broke[i] or np.max(resid_h[i]) <= self.tolerance),
                construction_wall_seconds=wall,
                action_calls=m,
                residual_tolerance=self.tolerance,
                criterion="gmres",
                peak_rss_kib=int(resource.getrusage(resource.RUSAGE_SELF).ru_maxrss),
            ))
        return out


def _drive(runs) -> None:
    """Advance every run from one host thread, one Krylov step of each in turn, until all stop."""

    handles = (ctypes.c_void_p * len(runs))(*[r.handle.value for r in runs])
    _lib.call("c3l_arnoldi_drive", handles, len(runs))


def _drive_and_finish(runs) -> list:
    """``_drive`` then ``finish`` every run; every run is ended even when driving fails."""

    try:
        _drive(runs)
    except BaseException:
        for r in runs:
            try:
                r.finish()
            except RuntimeError:
                pass
        raise
    return [r.finish() for r in runs]


def arnoldi_resolvent_batch(
    operators,
    rhs,
    shifts,
    *,
    maximum_dimension: int,
    residual_tolerance: float,
    sinks,
    minimum_dimension: int = 16,
    breakdown_tolerance: float = 1.0e-14,
) -> list[ArnoldiResolvent]:
    """``arnoldi_resolvent`` for several right-hand sides in lockstep (``c3l_arnoldi_batch``).

    ``operators[i]`` is the same sigma operator bound to distinct plan slots (twin plans);
    every sigma and orthogonalisation kernel of a step runs once over all bases that have
    not stopped.  Each basis stops on its own residual test.
    """

    run = _ArnoldiRun(operators, rhs, shifts, maximum_dimension=maximum_dimension,
                      residual_tolerance=residual_tolerance, sinks=sinks, minimum_dimension=minimum_dimension,
                      breakdown_tolerance=breakdown_tolerance)
    return _drive_and_finish([run])[0]


@dataclass(frozen=True)
class ComponentShiftedSolution:
    solution: np.ndarray
    converged: bool
    info: int
    iterations: int
    true_relative_residual: float
    residual_history: tuple
    wall_seconds: float


def solve_component_shifted(
    operator: ComponentSigmaOperator,
    shift: complex,
    rhs,
    *,
    rtol: float = 1.0e-10,
    restart: int = 40,
    maxiter: int = 200,
    initial_solution=None,
) -> ComponentShiftedSolution:
    """Solve ``(z I - σ) x = b`` on the GPU (``component_response.py:72-128``).

    The reference runs Jacobi-preconditioned restarted GMRES; here the correction
    ``x - x0`` is the GMRES iterate on one Arnoldi basis of σ built by ``c3l_arnoldi``
    (dimension cap ``min(P, restart*maxiter, 2048)``), recovered from the small
    least-squares problem on the Hessenberg matrix.  The true residual is re-evaluated
    with the full sigma action, and ``converged`` uses the reference's rule
    ``info == 0 and true <= max(10 rtol, 1e-12)``.
    """

    torch = _lib.require_device()
    n = int(operator.dimension)
    b = np.asarray(rhs, dtype=np.complex128)
    if b.shape != (n,) or not np.all(np.isfinite(b)):
        raise ValueError("shifted-response source has the wrong shape or is nonfinite")
    z = complex(shift)
    started = perf_counter()
    dev = _lib.DEVICE

    def apply(x: np.ndarray) -> np.ndarray:
        xd = torch.from_numpy(np.ascontiguousarray(x)).to(dev)
        out = torch.empty_like(xd)
        operator.matvec_device(xd, out)
        return z * x - out.cpu().numpy()

    x0 = np.zeros(n, dtype=np.complex128) if initial_solution is None else np.asarray(initial_solution, dtype=np.complex128)
    if x0.shape != (n,) or not np.all(np.isfinite(x0)):
        raise ValueError("initial_solution has the wrong shape or is nonfinite")
    r0 = b - apply(x0) if initial_solution is not None else b
    bnorm = max(float(np.linalg.norm(b)), np.finfo(float).tiny)
    x = x0.copy()
    history: tuple = ()
    iterations = 0
    info = 0
    if np.linalg.norm(r0) > 0.0:
        cap = min(n, restart * maxiter, 2048)
        proj = arnoldi_resolvent(operator, r0, np.array([z]), maximum_dimension=cap,
                                 residual_tolerance=rtol * bnorm / float(np.linalg.norm(r0)),
                                 minimum_dimension=min(16, cap), criterion="gmres")
        m = proj.dimension
        Hz = -proj.hessenberg[: m + 1, :m].copy()
        Hz[np.arange(m), np.arange(m)] += z
        rhs_small = np.zeros(m + 1, dtype=np.complex128)
        rhs_small[0] = proj.source_norm
        y = np.linalg.lstsq(Hz, rhs_small, rcond=None)[0]
        yd = torch.from_numpy(np.ascontiguousarray(y)).to(dev)
        corr = (yd @ proj.basis[:m]).cpu().numpy()
        x = x0 + corr
        iterations = m
        history = tuple(float(v) for v in proj.shift_relative_residuals)
        info = 0 if proj.converged else int(maxiter)
    true = float(np.linalg.norm(b - apply(x)) / bnorm)
    return ComponentShiftedSolution(
        solution=x,
        converged=bool(info == 0 and true <= max(10.0 * rtol, 1.0e-12)),
        info=int(info),
        iterations=int(iterations),
        true_relative_residual=true,
        residual_history=history,
        wall_seconds=float(perf_counter() - started),
    )


@dataclass
class ComponentRetardedGreen:
    omega: np.ndarray
    eta: float
    addition: np.ndarray
    removal: np.ndarray
    retarded: np.ndarray
    addition_projection: ArnoldiResolvent | None
    removal_projection: ArnoldiResolvent | None
    addition_relative_residuals: np.ndarray
    removal_relative_residuals: np.ndarray
    wall_seconds: float
    peak_rss_kib: int


def run_concurrently(tasks, streams: int | None = None):
    """Run independent device tasks on default-priority streams, one host thread per stream.

    With ``streams`` = k < len(tasks), stream i runs tasks i, i+k, ... one after the
    other.  The native drivers release the GIL, so the streams overlap on the GPU.
    Every stream first waits for the caller's stream; the caller's stream waits for
    all of them before returning.
    """

    torch = _lib.require_device()
    k = len(tasks) if streams is None else max(1, min(int(streams), len(tasks)))
    if k <= 1 or _lib.DEVICE != "cuda":
        return [task() for task in tasks]
    from concurrent.futures import ThreadPoolExecutor

    main = torch.cuda.current_stream()
    pool_streams = [torch.cuda.Stream() for _ in range(k)]
    for stream in pool_streams:
        stream.wait_stream(main)

    def run(i):
        with torch.cuda.stream(pool_streams[i]):
            return [(t, tasks[t]()) for t in range(i, len(tasks), k)]

    results = [None] * len(tasks)
    with ThreadPoolExecutor(max_workers=k) as pool:
        for group in pool.map(run, range(k)):
            for t, value in group:
                results[t] = value
    for stream in pool_streams:
        main.wait_stream(stream)
    return results


def _checked(proj, limit, name):
    if np.max(proj.shift_relative_residuals) > limit:
        raise RuntimeError(
            f"component EOM resolvent did not converge ({name}: max residual "
            f"{np.max(proj.shift_relative_residuals):.2e} at m={proj.dimension})"
        )
    return proj.green, proj, proj.shift_relative_residuals


def _channel(operator, source, sink, z, cap, rtol, limit, name):
    if not np.any(source):
        return np.zeros(z.size, dtype=np.complex128), None, np.zeros(z.size)
    proj = arnoldi_resolvent(operator, source, z, maximum_dimension=cap, residual_tolerance=rtol, sink=sink,
                             minimum_dimension=min(16, cap), criterion="gmres")
    return _checked(proj, limit, name)


def _block_runs(pairs, z, rtol, limit, maximum_dimension, streams):
    """Addition and removal resolvents of every block in lockstep on plan slots 0 and 1,
    block b on stream b % k, all driven from this host thread; results in pair order."""

    torch = _lib.require_device()
    main = torch.cuda.current_stream()
    k = len(pairs) if streams is None else max(1, min(int(streams), len(pairs)))
    pool = [torch.cuda.Stream() for _ in range(k)]
    results = [(np.zeros(z.size, dtype=np.complex128), None, np.zeros(z.size))] * (2 * len(pairs))
    runs, owners = [], []
    try:
        for b, (operator, sources) in enumerate(pairs):
            live = [(slot, name, src) for slot, (name, src) in
                    enumerate((("addition", sources.addition), ("removal", sources.removal))) if np.any(src)]
            if not live:
                continue
            cap = maximum_dimension or min(operator.dimension, 2048)
            stream = pool[b % k]
            stream.wait_stream(main)
            with torch.cuda.stream(stream):
                runs.append(_ArnoldiRun([operator.on_slot(slot) for slot, _, _ in live], [src for _, _, src in live],
                                        z, maximum_dimension=cap, residual_tolerance=rtol,
                                        sinks=[sources.sink] * len(live)))
            owners.append((b, live))
    except BaseException:
        for r in runs:
            try:
                r.finish()
            except RuntimeError:
                pass
        raise
    finished = _drive_and_finish(runs)
    for stream in pool:
        main.wait_stream(stream)
    for (b, live), projs in zip(owners, finished):
        for (slot, name, _), proj in zip(live, projs):
            results[2 * b + slot] = _checked(proj, limit, name)
    return results


def retarded_green_blocks(pairs, omega, eta: float, *, rtol: float = 1.0e-10, maximum_dimension: int | None = None,
                          concurrent: bool = True, streams: int | None = None,
                          batched: bool = True) -> list["ComponentRetardedGreen"]:
    """``component_retarded_green`` for several ``(operator, sources)`` charge blocks at once.

    With ``concurrent`` the addition and removal resolvents of a block use plan slots 0
    and 1 and the blocks run on concurrent streams (``streams`` of them, one per block or
    task by default): with ``batched`` both resolvents of a block advance in lockstep (one
    launch per kernel for both) and one host thread drives every block, step by step;
    otherwise every resolvent is a task of its own host thread.
    """

    started = perf_counter()
    frequency = np.asarray(omega, dtype=np.float64)
    z = frequency + 1j * float(eta)
    limit = max(10.0 * rtol, 1.0e-12)
    if concurrent and batched:
        results = _block_runs(pairs, z, rtol, limit, maximum_dimension, streams)
    else:
        tasks = []
        for operator, sources in pairs:
            cap = maximum_dimension or min(operator.dimension, 2048)
            for slot, (name, src) in enumerate((("addition", sources.addition), ("removal", sources.removal))):
                op = operator.on_slot(slot) if concurrent else operator
                tasks.append(lambda op=op, src=src, snk=sources.sink, cap=cap, name=name:
                             _channel(op, src, snk, z, cap, rtol, limit, name))
        results = run_concurrently(tasks, streams) if concurrent else [t() for t in tasks]
    out = []
    for b, (operator, _) in enumerate(pairs):
        (ga, pa, ra), (gr, pr, rr) = results[2 * b], results[2 * b + 1]
        if concurrent:  # slot operators counted their own matvecs
            operator.action_calls += sum(p.action_calls for p in (pa, pr) if p is not None)
        out.append(
            ComponentRetardedGreen(
                omega=frequency,
                eta=float(eta),
                addition=ga,
                removal=gr,
                retarded=ga + gr,
                addition_projection=pa,
                removal_projection=pr,
                addition_relative_residuals=ra,
                removal_relative_residuals=rr,
                wall_seconds=perf_counter() - started,
                peak_rss_kib=int(resource.getrusage(resource.RUSAGE_SELF).ru_maxrss),
            )
        )
    return out


def component_retarded_green(
    operator: ComponentSigmaOperator,
    sources: ComponentEOMSources,
    omega,
    eta: float,
    *,
    rtol: float = 1.0e-10,
    restart: int = 40,
    maxiter: int = 200,
    maximum_dimension: int | None = None,
    concurrent: bool = True,
) -> ComponentRetardedGreen:
    """Addition + removal contributions of ``G^R(ω)`` for one charge block.

    ``restart``/``maxiter`` are accepted for signature compatibility; the
    Krylov dimension is capped by ``maximum_dimension`` (default
    ``min(P, restart*maxiter, 2048)``).  Raises ``RuntimeError`` when a
    shift misses ``max(10 rtol, 1e-12)`` like the reference's shifted solves
    (``component_response.py:227-228``).
    """

    cap = maximum_dimension or min(operator.dimension, restart * maxiter, 2048)
    return retarded_green_blocks([(operator, sources)], omega, eta, rtol=rtol, maximum_dimension=cap,
                                 concurrent=concurrent)[0]


@dataclass(frozen=True)
class ComponentShiftedMatrixArnoldiGreen:
    """``G^R(z)[sink_i, source_j]`` over many complex shifts (axes shift, sink, source)
    (``component_response.py:449-465``)."""

    shifts: np.ndarray
    sites: tuple
    addition: np.ndarray
    removal: np.ndarray
    retarded: np.ndarray
    addition_projections: tuple
    removal_projections: tuple
    addition_relative_residuals: np.ndarray  # (source, shift) Galerkin residuals
    removal_relative_residuals: np.ndarray
    action_calls: int
    wall_seconds: float
    peak_rss_kib: int


@dataclass(frozen=True)
class ComponentRetardedMatrixArnoldiGreen:
    """Real-axis Green matrix over many frequencies (``component_response.py:162-178``)."""

    omega: np.ndarray
    eta: float
    sites: tuple
    addition: np.ndarray
    removal: np.ndarray
    retarded: np.ndarray
    addition_projections: tuple
    removal_projections: tuple
    addition_relative_residuals: np.ndarray
    removal_relative_residuals: np.ndarray
    component_action_calls: int
    component_action_wall_seconds: float
    wall_seconds: float
    peak_rss_kib: int


@dataclass(frozen=True)
class ComponentRetardedMatrixGreen:
    """Green matrix with one record per source, channel and frequency (``component_response.py:146-159``).

    The records are ``ComponentShiftedSolution`` objects whose ``solution`` is ``None`` (the
    vectors stay in the device Krylov space; ``*_projections`` reconstruct any of them) and whose
    ``true_relative_residual`` is the exact residual of that frequency's solution.
    """

    omega: np.ndarray
    eta: float
    sites: tuple
    addition: np.ndarray
    removal: np.ndarray
    retarded: np.ndarray
    addition_solutions: tuple
    removal_solutions: tuple
    addition_projections: tuple
    removal_projections: tuple
    wall_seconds: float
    peak_rss_kib: int


def _validated_source_columns(operator, sources):
    """Same checks and messages as ``component_response.py:496-517``; sources shorter than the
    operator (S13 sources in a larger space) are zero-padded."""

    columns = tuple(sources)
    if not columns:
        raise ValueError("at least one impurity source is required")
    if len({item.site for item in columns}) != len(columns):
        raise ValueError("impurity source sites must be unique")
    if any(item.charge != columns[0].charge for item in columns):
        raise ValueError("all impurity sources must use one response charge")
    retained = columns[0].addition.size
    if any(item.addition.shape != (retained,) or item.removal.shape != (retained,) or item.sink.shape != (retained,)
           for item in columns):
        raise ValueError("impurity sources and sinks must share one retained basis")
    if int(operator.dimension) < retained:
        raise ValueError("response operator is smaller than the retained sources")
    return columns, retained


def _validated_krylov_controls(operator, maximum_krylov_dimension, minimum_krylov_dimension, check_interval,
                               residual_tolerance):
    """``component_response.py:520-547``."""

    for value in (maximum_krylov_dimension, minimum_krylov_dimension, check_interval):
        if isinstance(value, bool) or not isinstance(value, (int, np.integer)):
            raise TypeError("Krylov dimensions and check_interval must be integers")
    if (maximum_krylov_dimension <= 0 or minimum_krylov_dimension <= 0 or check_interval <= 0
            or not np.isfinite(residual_tolerance) or residual_tolerance <= 0.0):
        raise ValueError("Krylov controls and residual tolerance must be positive")
    maximum = min(int(maximum_krylov_dimension), int(operator.dimension))
    return maximum, min(int(minimum_krylov_dimension), maximum)


def _pad(vector, padding):
    v = np.asarray(vector, dtype=np.complex128)
    return v if padding == 0 else np.concatenate((v, np.zeros(padding, dtype=np.complex128)))


def _matrix_arnoldi_response(operator, columns, retained, shifts, *, maximum_dimension, minimum_dimension,
                             residual_tolerance, check_interval) -> ComponentShiftedMatrixArnoldiGreen:
    """One Galerkin-stopped Krylov space per source and channel; every sink contracted from it
    (``component_response.py:550-625``)."""

    started = perf_counter()
    before = operator.action_calls
    padding = int(operator.dimension) - retained
    sinks = [_pad(item.sink, padding) for item in columns]
    shape = (shifts.size, len(columns), len(columns))
    values = {"addition": np.empty(shape, dtype=np.complex128), "removal": np.empty(shape, dtype=np.complex128)}
    projections = {"addition": [], "removal": []}
    residuals = {"addition": [], "removal": []}
    for j, column in enumerate(columns):
        for name in ("addition", "removal"):
            proj = arnoldi_resolvent(operator, _pad(getattr(column, name), padding), shifts,
                                     maximum_dimension=maximum_dimension, residual_tolerance=residual_tolerance,
                                     minimum_dimension=minimum_dimension, check_interval=check_interval,
                                     sink=sinks[0], criterion="galerkin")
            if not proj.converged:
                raise RuntimeError("a component EOM matrix Arnoldi projection failed")
            worst = np.zeros(shifts.size)
            for i, sink in enumerate(sinks):
                green, estimate = proj.project(sink, shifts)
                values[name][:, i, j] = green
                worst = np.maximum(worst, estimate)
            projections[name].append(proj)
            residuals[name].append(worst)
    expected = sum(p.dimension for p in projections["addition"] + projections["removal"])
    if operator.action_calls - before != expected:
        raise RuntimeError("component action count disagrees with Arnoldi dimensions")
    return ComponentShiftedMatrixArnoldiGreen(
        shifts=np.asarray(shifts, dtype=np.complex128),
        sites=tuple(item.site for item in columns),
        addition=values["addition"],
        removal=values["removal"],
        retarded=values["addition"] + values["removal"],
        addition_projections=tuple(projections["addition"]),
        removal_projections=tuple(projections["removal"]),
        addition_relative_residuals=np.asarray(residuals["addition"]),
        removal_relative_residuals=np.asarray(residuals["removal"]),
        action_calls=int(expected),
        wall_seconds=float(perf_counter() - started),
        peak_rss_kib=int(resource.getrusage(resource.RUSAGE_SELF).ru_maxrss),
    )


def component_retarded_matrix_green_shifts(
    operator: ComponentSigmaOperator,
    sources,
    shifts,
    *,
    maximum_krylov_dimension: int = 320,
    residual_tolerance: float = 1.0e-10,
    minimum_krylov_dimension: int = 16,
    check_interval: int = 8,
) -> ComponentShiftedMatrixArnoldiGreen:
    """Impurity Green matrix at complex shifts with ``Im z > 0``
    (``component_retarded_matrix_green_shifts``, ``component_response.py:449-493``).

    Imaginary-axis points ``z = i omega_n`` and real-axis points ``omega + i eta`` share one
    Krylov space per source and channel.  Each space is built by the native driver and stops
    when every shift's Galerkin residual meets ``residual_tolerance`` (the reference's rule);
    the value for every sink comes from the Schur form of the Hessenberg matrix (Galerkin
    coefficients, as ``ArnoldiResolvent.coefficients``, ``green/krylov.py:66-93``), with its
    residual ``|h_{m+1,m} e_m^T y| / beta`` returned per source, channel and shift.
    """

    columns, retained = _validated_source_columns(operator, sources)
    z = np.asarray(shifts, dtype=np.complex128)
    if z.ndim != 1 or z.size == 0 or not np.all(np.isfinite(z)) or np.any(z.imag <= 0.0):
        raise ValueError("shifts must be finite with positive imaginary parts")
    maximum, minimum = _validated_krylov_controls(operator, maximum_krylov_dimension, minimum_krylov_dimension,
                                                  check_interval, residual_tolerance)
    return _matrix_arnoldi_response(operator, columns, retained, z, maximum_dimension=maximum,
                                    minimum_dimension=minimum, residual_tolerance=float(residual_tolerance),
                                    check_interval=int(check_interval))


def _real_axis(omega, eta):
    frequency = np.asarray(omega, dtype=np.float64)
    broadening = float(eta)
    if (frequency.ndim != 1 or frequency.size == 0 or not np.all(np.isfinite(frequency))
            or not np.isfinite(broadening) or broadening <= 0.0):
        raise ValueError("omega and positive eta must be finite")
    return frequency, broadening


def component_retarded_matrix_green_arnoldi(
    operator: ComponentSigmaOperator,
    sources,
    omega,
    eta: float,
    *,
    maximum_krylov_dimension: int = 320,
    residual_tolerance: float = 1.0e-10,
    minimum_krylov_dimension: int = 16,
    check_interval: int = 8,
) -> ComponentRetardedMatrixArnoldiGreen:
    """Green matrix on ``omega + i eta`` from one Krylov space per source and channel
    (``component_response.py:356-446``; ``eta > 0`` as the reference requires)."""

    columns, retained = _validated_source_columns(operator, sources)
    frequency, broadening = _real_axis(omega, eta)
    maximum, minimum = _validated_krylov_controls(operator, maximum_krylov_dimension, minimum_krylov_dimension,
                                                  check_interval, residual_tolerance)
    before_seconds = operator.action_wall_seconds
    core = _matrix_arnoldi_response(operator, columns, retained, frequency.astype(np.complex128) + 1j * broadening,
                                    maximum_dimension=maximum, minimum_dimension=minimum,
                                    residual_tolerance=float(residual_tolerance), check_interval=int(check_interval))
    return ComponentRetardedMatrixArnoldiGreen(
        omega=frequency,
        eta=broadening,
        sites=core.sites,
        addition=core.addition,
        removal=core.removal,
        retarded=core.retarded,
        addition_projections=core.addition_projections,
        removal_projections=core.removal_projections,
        addition_relative_residuals=core.addition_relative_residuals,
        removal_relative_residuals=core.removal_relative_residuals,
        component_action_calls=core.action_calls,
        component_action_wall_seconds=float(operator.action_wall_seconds - before_seconds),
        wall_seconds=core.wall_seconds,
        peak_rss_kib=core.peak_rss_kib,
    )


def component_retarded_matrix_green(
    operator: ComponentSigmaOperator,
    sources,
    omega,
    eta: float,
    *,
    rtol: float = 1.0e-10,
    restart: int = 40,
    maxiter: int = 200,
) -> ComponentRetardedMatrixGreen:
    """Green matrix with every frequency's shifted systems solved to ``rtol``
    (``component_retarded_matrix_green``, ``component_response.py:248-353``).

    The reference runs one warm-started restarted GMRES per source, channel and frequency.
    Here one Krylov space per source and channel serves every frequency: it grows until the
    Galerkin residual of every frequency is at most ``rtol`` (the exact residual of that
    frequency's solution ``V_m y(z)``, by the Arnoldi relation), so every returned value
    meets the reference's acceptance ``true <= max(10 rtol, 1e-12)``; a space that cannot
    reach it within ``min(P, restart*maxiter, 2048)`` columns raises ``RuntimeError`` as the
    reference does for a failed shifted solve.
    """

    columns, retained = _validated_source_columns(operator, sources)
    frequency, broadening = _real_axis(omega, eta)
    if not (np.isfinite(rtol) and rtol > 0.0):
        raise ValueError("rtol must be finite and positive")
    cap = min(int(operator.dimension), int(restart) * int(maxiter), 2048)
    started = perf_counter()
    z = frequency.astype(np.complex128) + 1j * broadening
    core = _matrix_arnoldi_response(operator, columns, retained, z, maximum_dimension=cap,
                                    minimum_dimension=min(16, cap), residual_tolerance=float(rtol), check_interval=8)
    limit = max(10.0 * rtol, 1.0e-12)
    records = {}
    for name in ("addition", "removal"):
        worst = getattr(core, f"{name}_relative_residuals")
        if np.max(worst) > limit:
            raise RuntimeError("a component EOM matrix shifted solve failed")
        records[name] = tuple(
            tuple(ComponentShiftedSolution(None, True, 0, proj.dimension, float(worst[j, k]), (), 0.0)
                  for k in range(z.size))
            for j, proj in enumerate(getattr(core, f"{name}_projections"))
        )
    return ComponentRetardedMatrixGreen(
        omega=frequency,
        eta=broadening,
        sites=core.sites,
        addition=core.addition,
        removal=core.removal,
        retarded=core.retarded,
        addition_solutions=records["addition"],
        removal_solutions=records["removal"],
        addition_projections=core.addition_projections,
        removal_projections=core.removal_projections,
        wall_seconds=float(perf_counter() - started),
        peak_rss_kib=int(resource.getrusage(resource.RUSAGE_SELF).ru_maxrss),
    )


def component_retarded_time(
    operator: ComponentSigmaOperator,
    sources: ComponentEOMSources,
    time,
    *,
    maximum_krylov_dimension: int = 160,
    residual_tolerance: float = 1.0e-11,
) -> np.ndarray:
    """Positive-time impurity Green function ``G(t) = -i sink . exp(-i t sigma) (b_add + b_rem)``
    in one Arnoldi space (``component_retarded_time``, ``component_response.py:628-661``).

    The space is grown by the native driver with the reference's probe shift ``z = i`` and
    tolerances (Galerkin stop rule); ``exp(-i t H_m)`` acts on ``beta e_1`` on the host for each
    time, and the sink is projected on the basis on the device.
    """

    t = np.asarray(time, dtype=np.float64)
    if t.ndim != 1 or not np.all(np.isfinite(t)):
        raise ValueError("time must be a finite one-dimensional grid")
    p = np.asarray(sources.addition).size
    padding = int(operator.dimension) - p
    source = _pad(np.asarray(sources.addition, dtype=np.complex128) + np.asarray(sources.removal, dtype=np.complex128),
                  padding)
    sink = _pad(sources.sink, padding)
    n = int(operator.dimension)
    krylov = arnoldi_resolvent(operator, source, np.array([1j]), maximum_dimension=min(maximum_krylov_dimension, n),
                               residual_tolerance=residual_tolerance, minimum_dimension=min(16, n),
                               check_interval=min(8, n), sink=sink)
    if t.size == 0:
        return np.empty(0, dtype=np.complex128)
    m = krylov.dimension
    return -1j * (krylov.time_coefficients(t) @ krylov.projected_sink(sink, m))


@dataclass(frozen=True)
class ComponentRetardedMatrixTime:
    """Direct-time impurity Green matrix (``component_response.py:664-681``), axes (time, sink, source)."""

    time: np.ndarray
    sites: tuple
    addition: np.ndarray
    removal: np.ndarray
    retarded: np.ndarray
    channel_sum_defect: float
    combined_projections: tuple
    addition_projections: tuple
    removal_projections: tuple
    combined_relative_time_residuals: np.ndarray
    addition_relative_time_residuals: np.ndarray
    removal_relative_time_residuals: np.ndarray
    wall_seconds: float
    peak_rss_kib: int


def _matrix_time_from_projections(projections, sinks, time) -> np.ndarray:
    out = np.empty((time.size, len(sinks), len(projections)), dtype=np.complex128)
    for j, proj in enumerate(projections):
        m = proj.dimension
        projected = np.stack([proj.projected_sink(sink, m) for sink in sinks])  # (sink, m)
        out[:, :, j] = -1j * (proj.time_coefficients(time) @ projected.T)
    return out


def component_retarded_matrix_time(
    operator: ComponentSigmaOperator,
    sources,
    time,
    validation_shifts,
    *,
    maximum_krylov_dimension: int = 160,
    residual_tolerance: float = 1.0e-11,
    include_channels: bool = True,
) -> ComponentRetardedMatrixTime:
    """Positive-time Green matrix from one Krylov space per source orbital (and, with
    ``include_channels``, per channel) (``component_retarded_matrix_time``,
    ``component_response.py:702-815``).  Each space stops on the Galerkin residual at
    ``validation_shifts``; ``G_ab(t) = -i sink_a · V_m exp(-i t H_m) beta e_1``."""

    columns, retained = _validated_source_columns(operator, sources)
    t = np.asarray(time, dtype=np.float64)
    z = np.asarray(validation_shifts, dtype=np.complex128)
    if (t.ndim != 1 or t.size == 0 or np.any(t < 0.0) or not np.all(np.isfinite(t)) or z.ndim != 1 or z.size == 0
            or not np.all(np.isfinite(z))):
        raise ValueError("time and validation shifts must be finite nonempty vectors")
    padding = int(operator.dimension) - retained
    sinks = [_pad(item.sink, padding) for item in columns]
    n = int(operator.dimension)
    maximum = min(int(maximum_krylov_dimension), n)

    def project(vector):
        return arnoldi_resolvent(operator, _pad(vector, padding), z, maximum_dimension=maximum,
                                 residual_tolerance=residual_tolerance, minimum_dimension=min(16, n),
                                 check_interval=min(8, n), sink=sinks[0])

    started = perf_counter()
    combined = tuple(project(np.asarray(item.addition) + np.asarray(item.removal)) for item in columns)
    retarded = _matrix_time_from_projections(combined, sinks, t)
    combined_res = np.asarray([p.relative_time_residuals(t) for p in combined])
    if include_channels:
        add_p = tuple(project(item.addition) for item in columns)
        rem_p = tuple(project(item.removal) for item in columns)
        addition = _matrix_time_from_projections(add_p, sinks, t)
        removal = _matrix_time_from_projections(rem_p, sinks, t)
        add_res = np.asarray([p.relative_time_residuals(t) for p in add_p])
        rem_res = np.asarray([p.relative_time_residuals(t) for p in rem_p])
        defect = float(np.max(np.abs(retarded - addition - removal)))
    else:
        add_p = rem_p = ()
        add_res = np.empty((0, t.size))
        rem_res = np.empty((0, t.size))
        addition = np.full_like(retarded, np.nan + 1j * np.nan)
        removal = np.full_like(retarded, np.nan + 1j * np.nan)
        defect = float("nan")
    return ComponentRetardedMatrixTime(
        time=t,
        sites=tuple(item.site for item in columns),
        addition=addition,
        removal=removal,
        retarded=retarded,
        channel_sum_defect=defect,
        combined_projections=combined,
        addition_projections=add_p,
        removal_projections=rem_p,
        combined_relative_time_residuals=combined_res,
        addition_relative_time_residuals=add_res,
        removal_relative_time_residuals=rem_res,
        wall_seconds=float(perf_counter() - started),
        peak_rss_kib=int(resource.getrusage(resource.RUSAGE_SELF).ru_maxrss),
    )

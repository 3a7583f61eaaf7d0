"""Many independent model instances solved together: CCSD Newton--Krylov and the EOM Green
functions of a whole group of instances in lockstep on one GPU.

The reference solves instances one after the other (``solve_component_ccsd``,
``cc/component_solver.py:103-295``; ``component_retarded_green``, ``cc/component_response.py:180-245``;
timing driver ``examples/junction/gpu_timing.py:54-115``).  Small instances leave most of the GPU
idle, so here instances of equal site count form groups whose contraction plans are laid side by
side in one *union plan* (``plan_arrays.union_plan``): one launch per residual / JVP / sigma
program serves the whole group.  Every instance keeps its own Newton iteration (forcing terms,
line search, convergence test), its own GMRES (``c3l_gmres_group``: one union Jacobian product
per step, per-instance orthogonalisation and Givens state) and its own Arnoldi bases
(``c3l_arnoldi_group_begin``: addition and removal bases of every instance on two twin copies of
the union sigma plan), and drops out when it has converged.  Groups are independent and run one
after the other.

Instances whose default full-HF reference does not converge cannot be lowered (the reference
raises for them, ``cc/factorized.py:304-318``); ``prepare_batch`` leaves them out and lists them.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from time import perf_counter

import numpy as np

from . import _lib
from .device import DevicePlan, DeviceTilde, Reducer, raxpby, rscale
from .eom import ComponentEOMProblem, eom_ccsd_problem
from .lowering import TildeMap, factorized_cc_problem
from .plan_arrays import KINDS, PROGRAMS, union_plan
from .solver import NewtonKrylovPolicy, forcing_tolerance, rapidity_preconditioner
from .static import ComponentStaticProblem


def union_tilde(maps: list[TildeMap]) -> TildeMap:
    """Tilde maps of instances side by side (each instance's indices shifted into its segment)."""

    D = maps[0].dimension
    rd = maps[0].real_size
    if any(m.dimension != D or m.real_size != rd for m in maps):
        raise ValueError("united tilde maps must share their dimensions")
    shift = lambda a, k, size: np.concatenate([getattr(m, a) + i * size for i, m in enumerate(maps)])  # noqa: E731
    cat = lambda a: np.concatenate([getattr(m, a) for m in maps])  # noqa: E731
    return TildeMap(shift("partner", 0, D), cat("phase"), cat("alpha"), cat("beta"), shift("k0", 0, rd),
                    shift("k1", 0, rd), shift("pack_src", 0, D), cat("pack_w"))


@dataclass
class BatchInstance:
    index: int
    static: ComponentStaticProblem
    eom: dict  # charge -> ComponentEOMProblem


def compile_instance(index: int, model, charges=(1, -1)) -> BatchInstance:
    """Host set-up of one instance (full HF, lowering, plan compilation); no device work."""

    problem = factorized_cc_problem(model)
    static = ComponentStaticProblem.compile(problem)
    eom = eom_ccsd_problem(problem)
    return BatchInstance(index, static, {q: ComponentEOMProblem.compile(eom, static, q) for q in charges})


@dataclass
class InstanceGroup:
    """Instances of equal site count with their union plans on the device."""

    members: list  # BatchInstance
    charges: tuple
    site: int = 0
    tuning: list = field(default_factory=list)  # autotuner records of the union plans
    _dev: dict = field(default_factory=dict, repr=False)

    @property
    def count(self) -> int:
        return len(self.members)

    @property
    def D(self) -> int:
        if "D" not in self._dev:
            self._dev["D"] = self.members[0].static.oracle.num_amplitudes
        return self._dev["D"]

    @property
    def n(self) -> int:
        if "n" not in self._dev:
            self._dev["n"] = self.members[0].static.oracle.model.n_sites
        return self._dev["n"]

    def P(self, charge: int) -> int:
        if ("P", charge) not in self._dev:
            self._dev[("P", charge)] = self.members[0].eom[charge].dimension
        return self._dev[("P", charge)]

    def release_host(self) -> None:
        """Drop the host copies of the instances' problems and plans once the group is on the GPU
        (the device plans, maps and the small per-instance source plans stay); the group can then
        only be solved, not re-uploaded."""

        self.D, self.n  # noqa: B018  (cached before the problems go)
        for q in self.charges:
            self.P(q)
        for m in self.members:
            m.static.lowered = None
            m.static.oracle = None
            for e in m.eom.values():
                e.lowered = None
                e.static = None
                e.oracle = None
        for key, value in self._dev.items():
            for plan in (value if isinstance(value, list) else [value]):
                if isinstance(plan, DevicePlan):
                    plan.lowered = None

    def autotune(self, jvp_bits: int = 32, reps: int = 3) -> list:
        """Launch-shape autotuner (``tuning.py``) on the group's union plans: residual, JVP in
        ``jvp_bits`` and every charge's sigma; one tuning per union plan, at set-up."""

        torch = _lib.require_device()
        dev = _lib.DEVICE
        plan, diag = self._dev["plan"], self._dev["diag"]
        t = torch.zeros(self.count * self.D, dtype=torch.complex128, device=dev)
        out = torch.empty_like(t)
        plan.load("T", t)
        records = [plan.autotune("residual", out, diag, t, reps=reps)]
        if jvp_bits == 32:
            plan.set_precision("D", "jvp", 32)
        try:
            plan.load("D", t)
            records.append(plan.autotune("jvp", out, diag, t, reps=reps))
        finally:
            if jvp_bits == 32:
                plan.set_precision("D", "jvp", 64)
        for q in self.charges:
            sigma_plan = self._dev[("eom", q)][0]
            x = torch.zeros(self.count * self.P(q), dtype=torch.complex128, device=dev)
            sigma_plan.load("T", t)
            sigma_plan.load("X", x)
            records.append(dict(sigma_plan.autotune("sigma", torch.empty_like(x), self._dev[("eom_diag", q)], x,
                                                    reps=reps), charge=q))
        return records

    def upload(self) -> None:
        """Union plans (static; per charge two twin copies of the sigma plan), union tilde maps,
        per-instance source plans and sinks, all resident on the GPU."""

        torch = _lib.require_device()
        dev = _lib.DEVICE
        G, D = self.count, self.D
        statics = [m.static for m in self.members]
        static_plan = union_plan([s.lowered for s in statics], {KINDS["T"]: D, KINDS["D"]: D},
                                 {PROGRAMS["residual"]: D, PROGRAMS["reference"]: 1, PROGRAMS["jvp"]: D,
                                  PROGRAMS["residual_low"]: D})
        self._dev["plan"] = DevicePlan(static_plan)
        self._dev["tilde"] = DeviceTilde(union_tilde([s.oracle.tilde for s in statics]))
        self._dev["diag"] = torch.from_numpy(np.concatenate([s.diagonal for s in statics])).to(dev)
        self._dev["inv_diag"] = torch.from_numpy(np.concatenate([rapidity_preconditioner(s)[0] for s in statics])).to(dev)
        for q in self.charges:
            P = self.P(q)
            lowered = union_plan([m.eom[q].lowered for m in self.members], {KINDS["T"]: D, KINDS["X"]: P},
                                 {PROGRAMS["sigma"]: P})
            self._dev[("eom", q)] = [DevicePlan(lowered), DevicePlan(lowered)]  # twins: addition, removal
            self._dev[("eom_diag", q)] = torch.from_numpy(
                np.concatenate([m.eom[q].diagonal for m in self.members])).to(dev)
            for m in self.members:
                m.eom[q].source_device(self.site)
            self._dev[("sink", q)] = np.stack([m.eom[q].sink(self.site) for m in self.members])


@dataclass
class BatchSolution:
    index: int
    converged: bool
    packed_residual_norm: float
    newton_iterations: int
    jvp_evaluations: int
    residual_evaluations: int
    amplitudes: np.ndarray
    green: dict = field(default_factory=dict)  # charge -> G^R on the grid
    krylov_dimensions: dict = field(default_factory=dict)
    max_shift_residual: float = 0.0


def group_newton(group: InstanceGroup, policy: NewtonKrylovPolicy | None = None) -> list[dict]:
    """Lockstep inexact Newton--Krylov for every instance of the group (``solve._newton`` per
    instance: the same forcing terms, GMRES, acceptance test, Armijo line search and final packed
    residual), with one union launch per residual, JVP and orthogonalisation step for all."""

    torch = _lib.require_device()
    controls = policy or NewtonKrylovPolicy.mixed_precision(forcing_strategy="eisenstat_walker", gmres_restart=150)
    dev = _lib.DEVICE
    G, D = group.count, group.D
    plan, tilde = group._dev["plan"], group._dev["tilde"]
    diag, inv_diag = group._dev["diag"], group._dev["inv_diag"]
    f64, c128 = torch.float64, torch.complex128
    N = G * D
    x = torch.zeros(N, dtype=f64, device=dev)
    r = torch.empty(N, dtype=f64, device=dev)
    trial_r = torch.empty(N, dtype=f64, device=dev)
    step = torch.empty(N, dtype=f64, device=dev)
    xtrial = torch.empty(N, dtype=f64, device=dev)
    b = torch.empty(N, dtype=f64, device=dev)
    w = torch.empty(N, dtype=f64, device=dev)
    tmp = torch.empty(N, dtype=f64, device=dev)
    t = torch.zeros(N, dtype=c128, device=dev)
    dt = torch.zeros(N, dtype=c128, device=dev)
    out = torch.zeros(N, dtype=c128, device=dev)
    ref = torch.zeros(G, dtype=c128, device=dev)
    R = controls.gmres_restart
    bits = 32 if controls.jacobian_precision == "fp32" else 64
    V = torch.zeros(R + 1, N, dtype=torch.float32 if bits == 32 else f64, device=dev)
    scratch_size, coef_size = _lib.krylov_workspace(False, D, R)
    scratch = torch.zeros(G * scratch_size, dtype=f64, device=dev)
    coef = torch.zeros(G * coef_size, dtype=f64, device=dev)
    red = Reducer(D, 1, complex_=False)
    sq = torch.empty(G, dtype=f64, device=dev)
    maps = _lib.TildeMaps(tilde.dim, tilde.real_dim, _lib.dptr(tilde.k0), _lib.dptr(tilde.k1), _lib.dptr(tilde.ab),
                          _lib.dptr(tilde.src), _lib.dptr(tilde.w), _lib.dptr(inv_diag))
    seg = [slice(i * D, (i + 1) * D) for i in range(G)]
    stats = [dict(residual_calls=0, jvp_calls=0, iterations=0, converged=False, failed=False) for _ in range(G)]

    def residual(xv, rv) -> np.ndarray:
        tilde.unpack(xv, t)
        plan.load("T", t)
        plan.apply("residual", out, diag, t)
        plan.apply("reference", ref)
        tilde.pack(out, rv)
        for i in range(G):
            red.dot(rv[seg[i]].view(1, -1), 1, rv[seg[i]], sq[i:i + 1])
        return np.sqrt(np.maximum(sq.cpu().numpy(), 0.0))

    def dots(a, c) -> np.ndarray:
        for i in range(G):
            red.dot(a[seg[i]].view(1, -1), 1, c[seg[i]], sq[i:i + 1])
        return sq.cpu().numpy().copy()

    plan.set_precision("D", "jvp", bits)
    try:
        norm = residual(x, r)
        for i in range(G):
            stats[i]["residual_calls"] += 1
        live = [i for i in range(G)]
        previous = [None] * G
        previous_tol = [None] * G
        for iteration in range(controls.maximum_newton_iterations + 1):
            for i in list(live):
                if not np.isfinite(norm[i]):
                    stats[i]["failed"] = True
                    live.remove(i)
                elif norm[i] <= controls.tolerance:
                    stats[i]["converged"] = True
                    live.remove(i)
            if not live or iteration == controls.maximum_newton_iterations:
                break
            tol = np.ones(G)
            for i in live:
                tol[i] = forcing_tolerance(norm[i], controls, previous[i], previous_tol[i])
            atol = np.zeros(G) if controls.forcing_strategy == "legacy" else np.full(
                G, controls.gmres_absolute_tolerance_factor * controls.tolerance)
            rscale(r, b, -1.0, torch)
            active = (ctypes.c_int32 * G)(*[1 if i in live else 0 for i in range(G)])
            rt = (ctypes.c_double * G)(*tol.tolist())
            at = (ctypes.c_double * G)(*atol.tolist())
            info, its, calls = (ctypes.c_int32 * G)(), (ctypes.c_int32 * G)(), (ctypes.c_int32 * G)()
            true = (ctypes.c_double * G)()
            _lib.call("c3l_gmres_group", plan.handle, ctypes.byref(maps), G, _lib.dptr(diag), _lib.dptr(dt),
                      _lib.dptr(out), _lib.dptr(b), _lib.dptr(step), _lib.dptr(V), bits, _lib.dptr(w), _lib.dptr(tmp),
                      _lib.dptr(scratch), _lib.dptr(coef), rt, at, active, R, controls.gmres_maximum_restart_cycles,
                      info, its, true, calls, _lib.stream_ptr(torch))
            for i in live:
                stats[i]["jvp_calls"] += calls[i]
            if controls.forcing_strategy == "eisenstat_walker":
                rw = dots(r, w)
                for i in list(live):
                    slope = -norm[i] ** 2 - rw[i]
                    defect = true[i] * norm[i]
                    target = max(tol[i] * norm[i], atol[i])
                    if not (np.isfinite(defect) and defect <= (1.0 + np.sqrt(np.finfo(float).eps)) * target
                            and np.isfinite(slope) and slope < 0.0):
                        stats[i]["failed"] = True  # the single-instance solver stops here too
                        live.remove(i)
            damping = {i: 1.0 for i in live}
            pending = list(live)
            accepted = {}
            xtrial.copy_(x)
            for _ in range(controls.maximum_line_search_steps):
                for i in pending:
                    xtrial[seg[i]].copy_(x[seg[i]])
                    raxpby(damping[i], step[seg[i]], 1.0, xtrial[seg[i]], torch)
                trial_norm = residual(xtrial, trial_r)
                for i in range(G):
                    if i in pending:
                        stats[i]["residual_calls"] += 1
                for i in list(pending):
                    if trial_norm[i] <= (1.0 - controls.armijo * damping[i]) * norm[i]:
                        x[seg[i]].copy_(xtrial[seg[i]])
                        r[seg[i]].copy_(trial_r[seg[i]])
                        accepted[i] = trial_norm[i]
                        pending.remove(i)
                    else:
                        damping[i] *= 0.5
                if not pending:
                    break
            for i in pending:
                stats[i]["failed"] = True
                live.remove(i)
            for i, value in accepted.items():
                if i in live:
                    previous[i], previous_tol[i] = norm[i], tol[i]
                    norm[i] = value
                    stats[i]["iterations"] += 1
    finally:
        if bits != 64:
            plan.set_precision("D", "jvp", 64)
    packed = residual(x, r)  # t and out now belong to x
    for i in range(G):
        stats[i]["residual_calls"] += 1
        stats[i]["packed_residual_norm"] = float(packed[i])
        stats[i]["converged"] = bool(stats[i]["converged"] and packed[i] <= controls.tolerance)
    group._dev["t"] = t
    return stats


def group_green(group: InstanceGroup, omega, eta: float = 0.0, rtol: float = 1.0e-10,
                maximum_dimension: int | None = None) -> dict:
    """Addition and removal resolvents of every instance and charge on the group's amplitudes
    (``group._dev["t"]`` from ``group_newton``): per charge one group run of 2 x G bases on the
    two twin copies of the union sigma plan, both charges driven by one host thread."""

    torch = _lib.require_device()
    dev = _lib.DEVICE
    G, D = group.count, group.D
    t = group._dev["t"]
    z = np.ascontiguousarray(np.asarray(omega, dtype=np.float64) + 1j * float(eta))
    ns = z.size
    zs = torch.from_numpy(z).to(dev)
    main = torch.cuda.current_stream()
    runs, keep, meta = [], [], []
    try:
        for q in group.charges:
            P = group.P(q)
            slots = group._dev[("eom", q)]
            m_max = int(maximum_dimension or min(P, 1024))
            stream = torch.cuda.Stream()
            stream.wait_stream(main)
            with torch.cuda.stream(stream):
                for s in slots:
                    s.load("T", t)
                add = torch.empty(G, P, dtype=torch.complex128, device=dev)
                rem = torch.empty(G, P, dtype=torch.complex128, device=dev)
                for i, m in enumerate(group.members):
                    src = m.eom[q].source_device(group.site)
                    src.load("T", t[i * D:(i + 1) * D])
                    src.apply("source_addition", add[i])
                    src.apply("source_removal", rem[i])
                Bh = np.stack([add.cpu().numpy(), rem.cpu().numpy()])  # [slot][instance][P]
                beta = np.linalg.norm(Bh, axis=2)
                if np.any(beta == 0.0):
                    raise ValueError("every instance needs nonzero addition and removal sources")
                count = 2 * G
                V = torch.empty(2, m_max + 1, G * P, dtype=torch.complex128, device=dev)
                V[:, 0].copy_(torch.from_numpy((Bh / beta[:, :, None]).reshape(2, G * P)))
                beta = beta.reshape(-1)
                W = torch.empty(2, G * P, dtype=torch.complex128, device=dev)
                H = torch.zeros(count, m_max, m_max + 1, dtype=torch.complex128, device=dev)
                sink = torch.from_numpy(np.ascontiguousarray(np.concatenate([group._dev[("sink", q)]] * 2))).to(dev)
                rot_c = torch.empty(count, ns * m_max, dtype=torch.float64, device=dev)
                rot_s = torch.empty(count, ns * m_max, dtype=torch.complex128, device=dev)
                pi = torch.empty(count, ns * m_max, dtype=torch.complex128, device=dev)
                gamma = torch.zeros(count, ns, dtype=torch.complex128, device=dev)
                green = torch.zeros(count, ns, dtype=torch.complex128, device=dev)
                resid = torch.zeros(count, ns, dtype=torch.float64, device=dev)
                scratch_size, coef_size = _lib.krylov_workspace(True, P, m_max, ns)
                scratch = torch.empty(count, scratch_size, dtype=torch.float64, device=dev)
                coef = torch.zeros(count, coef_size, dtype=torch.float64, device=dev)
                status = torch.zeros(count, 4, dtype=torch.float64, device=dev)
                plans = (ctypes.c_void_p * 2)(*[s.handle.value for s in slots])
                beta_h = (ctypes.c_double * count)(*beta.tolist())
                handle = ctypes.c_void_p()
                _lib.call("c3l_arnoldi_group_begin", plans, 2, G, _lib.dptr(group._dev[("eom_diag", q)]), _lib.dptr(V),
                          _lib.dptr(W), _lib.dptr(H), P, m_max, _lib.dptr(zs), ns, _lib.dptr(sink), beta_h,
                          _lib.dptr(rot_c), _lib.dptr(rot_s), _lib.dptr(pi), _lib.dptr(gamma), _lib.dptr(green),
                          _lib.dptr(resid), _lib.dptr(scratch), _lib.dptr(coef), _lib.dptr(status), float(rtol), 16,
                          1.0e-14, _lib.stream_ptr(torch), ctypes.byref(handle))
            runs.append(handle)
            keep.append((V, W, H, sink, rot_c, rot_s, pi, gamma, scratch, coef, status, add, rem, stream))
            meta.append((q, green, resid, count))
        handles = (ctypes.c_void_p * len(runs))(*[h.value for h in runs])
        _lib.call("c3l_arnoldi_drive", handles, len(runs))
    finally:
        finished = []
        for h, (q, green, resid, count) in zip(runs, meta):
            m_out, broke, worst = (ctypes.c_int32 * count)(), (ctypes.c_int32 * count)(), (ctypes.c_double * count)()
            _lib.call("c3l_arnoldi_end", h, m_out, broke, worst)
            finished.append((q, green, resid, list(m_out)))
        for item in keep:
            main.wait_stream(item[-1])
    out = {}
    for q, green, resid, dims in finished:
        g = green.cpu().numpy().reshape(2, G, ns)
        res = resid.cpu().numpy().reshape(2, G, ns)
        out[q] = dict(addition=g[0], removal=g[1], retarded=g[0] + g[1], residuals=res,
                      dimensions=np.asarray(dims).reshape(2, G))
    return out


def group_by_size(instances: list[BatchInstance], max_group: int = 2) -> list[list[BatchInstance]]:
    """Instances of equal site count in groups of at most ``max_group`` (in instance order).
    Two per group measured fastest over the 512-instance uneven batch on MI355X (11.2 solves/s
    aggregate against 10.8 with four and 9.7 with eight per group, same box)."""

    by_n: dict[int, list] = {}
    for inst in instances:
        by_n.setdefault(inst.static.oracle.model.n_sites, []).append(inst)
    groups = []
    for n in sorted(by_n):
        members = by_n[n]
        k = -(-len(members) // max_group)  # as even as possible
        size = -(-len(members) // k)
        groups.extend(members[s:s + size] for s in range(0, len(members), size))
    return groups


def solve_group(group: InstanceGroup, omega, eta: float = 0.0, policy=None, rtol: float = 1.0e-10) -> list[BatchSolution]:
    """One complete solve (CCSD + both charges' G^R) of every instance of a group."""

    if "plan" not in group._dev:
        group.upload()
    stats = group_newton(group, policy)
    if not all(s["converged"] for s in stats):
        bad = [m.index for m, s in zip(group.members, stats) if not s["converged"]]
        raise RuntimeError(f"CCSD did not converge for instances {bad}")
    green = group_green(group, omega, eta, rtol)
    t = group._dev["t"].cpu().numpy()
    D = group.D
    limit = max(10.0 * rtol, 1.0e-12)
    out = []
    for i, (m, s) in enumerate(zip(group.members, stats)):
        worst = max(float(np.max(green[q]["residuals"][:, i])) for q in green)
        if worst > limit:
            raise RuntimeError(f"component EOM resolvent did not converge for instance {m.index} ({worst:.2e})")
        out.append(BatchSolution(
            index=m.index, converged=s["converged"], packed_residual_norm=s["packed_residual_norm"],
            newton_iterations=s["iterations"], jvp_evaluations=s["jvp_calls"], residual_evaluations=s["residual_calls"],
            amplitudes=t[i * D:(i + 1) * D].copy(),
            green={q: green[q]["retarded"][i].copy() for q in green},
            krylov_dimensions={q: tuple(int(d) for d in green[q]["dimensions"][:, i]) for q in green},
            max_shift_residual=worst,
        ))
    return out


def _compile_job(args):
    index, model, charges = args[:3]
    try:
        if len(args) > 3 and args[3]:
            # worker process: the host set-up is many small dense factorisations (Sylvester
            # solves of 2n x 2n), which run fastest on one BLAS thread per worker
            from threadpoolctl import threadpool_limits

            with threadpool_limits(limits=1):
                return compile_instance(index, model, charges), None
        return compile_instance(index, model, charges), None
    except RuntimeError as err:
        if "full-HF" not in str(err):
            raise
        return None, (index, str(err))


@dataclass
class PreparedBatch:
    groups: list  # InstanceGroup
    skipped: list  # (instance index, reason)
    setup_seconds: float

    @property
    def instances(self) -> int:
        return sum(g.count for g in self.groups)


def prepare_batch(models, charges=(1, -1), *, max_group: int = 2, workers: int = 1, indices=None,
                  release_host: bool = False, autotune: bool = True, jvp_bits: int = 32) -> PreparedBatch:
    """Host set-up of every instance (in ``workers`` processes when > 1; no GPU is touched there),
    grouping by site count, upload of the groups' union plans and (``autotune``) the launch-shape
    autotuner on them, with the JVP in ``jvp_bits``.  Instances whose full-HF reference does not
    converge are skipped and listed.  ``indices`` names the instances (default 0..)."""

    started = perf_counter()
    names = range(len(models)) if indices is None else indices
    jobs = [(int(i), m, tuple(charges), workers > 1) for i, m in zip(names, models)]
    if workers > 1:
        import multiprocessing as mp
        from concurrent.futures import ProcessPoolExecutor

        with ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as pool:
            results = list(pool.map(_compile_job, jobs, chunksize=1))
    else:
        results = [_compile_job(j) for j in jobs]
    instances = [inst for inst, _ in results if inst is not None]
    skipped = [why for _, why in results if why is not None]
    del results
    groups = [InstanceGroup(members, tuple(charges)) for members in group_by_size(instances, max_group)]
    for g in groups:
        g.upload()
        if autotune:
            g.tuning = g.autotune(jvp_bits)
        if release_host:
            g.release_host()
    return PreparedBatch(groups, skipped, perf_counter() - started)

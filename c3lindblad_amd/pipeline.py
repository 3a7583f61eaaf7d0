"""One complete CCSD + EOM G^R(ω) solve of a model: the unit of the solves/sec metric.

``prepare`` does the host set-up once (full HF, normal modes, lowering, plan
compilation for the static problem and each requested charge block).
``SolvePipeline.solve`` is one solve on the GPU: Newton from the zero guess to
``||pack R|| <= tolerance``, an EOM context per charge, and the addition and
removal resolvents on every frequency.  Nothing computed by one solve is
reused by the next apart from the compiled plans.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from time import perf_counter

import numpy as np

from .eom import ComponentEOMProblem, ComponentSigmaOperator, component_eom_context, component_eom_sources, eom_ccsd_problem
from .lowering import factorized_cc_problem
from .response import retarded_green_blocks
from .solver import NewtonKrylovPolicy, solve_component_ccsd
from .static import ComponentStaticProblem


@dataclass
class SolveResult:
    green: dict  # charge -> complex G^R on the grid
    packed_residual_norm: float
    newton_iterations: int
    jvp_evaluations: int
    residual_evaluations: int
    krylov_dimensions: dict
    max_shift_residual: float
    seconds: dict = field(default_factory=dict)
    amplitudes: np.ndarray | None = None
    gmres_iterations: tuple = ()  # per Newton step


@dataclass
class SolvePipeline:
    static: ComponentStaticProblem
    eom: dict  # charge -> ComponentEOMProblem
    omega: np.ndarray
    eta: float
    setup_seconds: float
    policy: NewtonKrylovPolicy
    rtol: float = 1e-10
    concurrent: bool = True
    eom_streams: int | None = None  # streams for the EOM resolvent tasks (None: one per task)
    tuning: list = field(default_factory=list)  # autotuner records of the set-up

    @classmethod
    def prepare(cls, model, omega, eta=0.0, charges=(1, -1), policy=None, rtol=1e-10, autotune=True, space="S13"):
        """Host set-up and upload; ``autotune`` runs the launch-shape autotuner on the residual,
        JVP (in the policy's precision) and sigma programs (``tuning.py``), inside the set-up time.
        ``space="S135"`` solves the response on S1 + S3 + the rank-five shell."""

        started = perf_counter()
        problem = factorized_cc_problem(model)
        static = ComponentStaticProblem.compile(problem)
        eom = eom_ccsd_problem(problem, space=space)
        blocks = {q: ComponentEOMProblem.compile(eom, static, q) for q in charges}
        static.device  # upload
        for b in blocks.values():
            b.device_plan(0)
            b.device_plan(1)
            b.source_device(model.impurity_site)
            b.sink(model.impurity_site)
        policy = policy or NewtonKrylovPolicy()
        tuning = []
        if autotune:
            from .tuning import autotune_sigma, autotune_static

            tuning = autotune_static(static, 32 if policy.jacobian_precision == "fp32" else 64)
            tuning += [dict(autotune_sigma(b), charge=q) for q, b in blocks.items()]
        return cls(static, blocks, np.asarray(omega, dtype=np.float64), float(eta), perf_counter() - started,
                   policy, rtol, tuning=tuning)

    def solve(self) -> SolveResult:
        t0 = perf_counter()
        sol = solve_component_ccsd(self.static, policy=self.policy)
        if not sol.converged:
            raise RuntimeError(f"CCSD did not converge (packed residual {sol.packed_residual_norm:.2e})")
        t1 = perf_counter()
        green = {}
        dims = {}
        worst = 0.0
        charges = list(self.eom)
        pairs = []
        for q in charges:
            ctx = component_eom_context(self.eom[q], sol.amplitudes, device_amplitudes=sol.device_amplitudes)
            pairs.append((ComponentSigmaOperator(ctx), component_eom_sources(ctx)))
        t_sources = perf_counter()
        results = retarded_green_blocks(pairs, self.omega, self.eta, rtol=self.rtol, concurrent=self.concurrent,
                                        streams=self.eom_streams)
        for q, result in zip(charges, results):
            green[q] = result.retarded
            dims[q] = tuple(p.dimension if p is not None else 0 for p in (result.addition_projection, result.removal_projection))
            worst = max(worst, float(np.max(result.addition_relative_residuals)), float(np.max(result.removal_relative_residuals)))
        t2 = perf_counter()
        return SolveResult(
            green=green,
            packed_residual_norm=sol.packed_residual_norm,
            newton_iterations=sol.iterations,
            jvp_evaluations=sol.jvp_evaluations,
            residual_evaluations=sol.residual_evaluations,
            krylov_dimensions=dims,
            max_shift_residual=worst,
            seconds={"ccsd": t1 - t0, "eom_green": t2 - t1, "eom_contexts_sources": t_sources - t1},
            amplitudes=sol.amplitudes,
            gmres_iterations=tuple(step.gmres_iterations for step in sol.history),
        )


def solve_concurrently(pipelines, rounds: int = 1) -> tuple[list, float]:
    """Run every pipeline ``rounds`` times, all pipelines at once, one host thread and
    one default-priority stream each; return the results (per pipeline, per round) and
    the wall time.

    Independent instances are latency bound (Krylov steps, small contractions), so
    interleaving several of them on one GPU raises the aggregate solves per second.
    """

    from concurrent.futures import ThreadPoolExecutor

    from . import _lib

    torch = _lib.require_device()
    main = torch.cuda.current_stream()
    streams = [torch.cuda.Stream() for _ in pipelines]
    for stream in streams:
        stream.wait_stream(main)

    def run(i):
        out = []
        with torch.cuda.stream(streams[i]):
            for _ in range(rounds):
                out.append(pipelines[i].solve())
        streams[i].synchronize()
        return out

    torch.cuda.synchronize()
    started = perf_counter()
    with ThreadPoolExecutor(max_workers=len(pipelines)) as pool:
        results = list(pool.map(run, range(len(pipelines))))
    torch.cuda.synchronize()
    elapsed = perf_counter() - started
    for stream in streams:
        main.wait_stream(stream)
    return results, elapsed


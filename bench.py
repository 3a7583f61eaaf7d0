"""Benchmark: CCSD + EOM G^R(ω) solves/sec on one MI355X.

Workload (default ``siam24``): single-impurity Anderson model with 24 Lindblad
pseudomodes (n = 25 sites, U = 4, eps_d = -2, seeded), CCSD steady state to
``||pack R|| <= 1e-10``, then charge +1 and -1 EOM-CCSD (S1+S3) retarded Green
functions on 4001 frequencies in [-8, 8] with eta = 0.  One step is one full
solve from the zero initial guess; host set-up (full HF, normal modes, plan
compilation and upload) happens once before timing and is reported as
``config.setup_seconds``.  ``--config readme`` runs the 3-site README model.

Default variant: mixed precision (FP32 JVPs inside GMRES; FP64 residual, update,
orthogonalisation, sigma and resolvent), Eisenstat-Walker forcing, GMRES restart
150.  After the timed steps one solve with the reference's default policy (FP64,
legacy forcing, restart 40) is timed as well and the G^R difference between the
two is reported (``config.fp64_reference_policy``).

Prints one JSON line (rank 0).  ``roofline`` is for the kernel family with the
largest total time in one extra instrumented solve after the timed steps (EOM
channels serial, every plan program and every driver orthogonalisation call
bracketed by HIP events on its stream); ``cpu_baseline`` times the NumPy CPU port
of the same plans on a bounded sample (see ``sample``).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0


def _lib_handle():
    from c3lindblad_amd import _lib

    return _lib.load_library()


def _measured_traffic(kernel: str):
    """HBM bytes per launch of ``kernel`` from the committed counter summaries in ``profiles/``
    (lines ``traffic_bytes_per_launch[<kernel>] = <bytes>``, FETCH_SIZE + WRITE_SIZE scaled by a
    copy of known size), or ``(None, None)``."""

    import glob
    import re

    here = os.path.dirname(os.path.abspath(__file__))
    pattern = re.compile(r"^traffic_bytes_per_launch\[(.+)\] = ([0-9.eE+]+)\s*$")
    for path in sorted(glob.glob(os.path.join(here, "profiles", "*.md"))):
        with open(path) as f:
            for line in f:
                m = pattern.match(line.strip())
                if m and m.group(1) == kernel:
                    return float(m.group(2)), os.path.relpath(path, here)
    return None, None


def _algorithmic_program_bytes(lowered, program: int, atom_bytes: int = 16) -> float:
    """Compulsory bytes of one program launch: per coordinate its output (complex128),
    destination, packed coordinate and (residual/JVP/sigma) diagonal and input value, plus
    every distinct atom tensor the program reads once, at ``atom_bytes`` per element (16 for
    FP64 atoms, 8 for the complex64 atoms of an FP32 program)."""

    n = lowered.n
    total = 0.0
    atoms = {}
    for prog, deg, coord, dest, coeff, fstart, foff, fcode in lowered.blocks:
        if prog != program:
            continue
        diag = 32 if program in (0, 2, 3) else 0
        total += dest.size * (16 + 8 + 4 + diag)
        for off, code in zip(foff, fcode):
            atoms[int(off)] = n ** (int(code) & 7)
    return total + float(atom_bytes) * sum(atoms.values())


def _plain_fp64_estimate(pipe, ref, runs=((0, 4), (1000, 4), (2000, 4), (3000, 4))) -> dict:
    """Throughput of the plain FP64 GPU path -- the reference's algorithm for the Green
    function, one shifted GMRES solve per frequency and channel warm-started from the
    previous frequency (``component_response.py:180-245``), with no basis shared between
    frequencies -- estimated from a timed sample: consecutive-frequency runs ``(start,
    length)`` of every channel, each solve on this GPU (``solve_component_shifted``), mean
    time per frequency scaled to the whole grid, plus the measured FP64 CCSD time of
    ``ref``.  The sampled G values are also compared with the shared-basis result."""

    import torch

    from c3lindblad_amd.eom import ComponentSigmaOperator, component_eom_context, component_eom_sources
    from c3lindblad_amd.response import solve_component_shifted

    omega = pipe.omega
    per_channel = {}
    worst_gap = 0.0
    solves = 0
    for q, problem in pipe.eom.items():
        ctx = component_eom_context(problem, ref.amplitudes)
        op = ComponentSigmaOperator(ctx)
        sources = component_eom_sources(ctx)
        sample = np.zeros(0, dtype=np.complex128)
        for name, src in (("addition", sources.addition), ("removal", sources.removal)):
            if not np.any(src):
                continue
            seconds = 0.0
            count = 0
            green = []
            for start, length in runs:
                x = None
                for k in range(start, min(start + length, omega.size)):
                    torch.cuda.synchronize()
                    t0 = time.perf_counter()
                    sol = solve_component_shifted(op, complex(omega[k] + 1j * pipe.eta), src, rtol=pipe.rtol,
                                                  initial_solution=x)
                    seconds += time.perf_counter() - t0
                    count += 1
                    x = sol.solution
                    green.append(complex(np.dot(sources.sink, x)))
            per_channel[f"{q}:{name}"] = seconds / count
            solves += count
            sample = np.asarray(green) if sample.size == 0 else sample + np.asarray(green)
        idx = np.concatenate([np.arange(s, min(s + n, omega.size)) for s, n in runs])
        full = ref.green[q][idx]
        worst_gap = max(worst_gap, float(np.max(np.abs(sample - full)) / np.max(np.abs(ref.green[q]))))
    total = ref.seconds["ccsd"] + sum(per_channel.values()) * omega.size
    return {
        "method": "FP64 CCSD (reference policy) + one warm-started shifted solve per frequency and channel",
        "solves_per_sec": 1.0 / total,
        "ccsd_seconds": ref.seconds["ccsd"],
        "seconds_per_frequency": {k: round(v, 6) for k, v in per_channel.items()},
        "sample": f"{solves} timed shifted solves in runs of consecutive frequencies starting at grid indices "
                  f"{[s for s, _ in runs]}, scaled to {omega.size} frequencies",
        "sampled_gr_rel_gap_to_shared_basis": worst_gap,
    }


def _host_cgs_seconds(rows: int, columns: list, dtype, threads: int) -> float:
    """Host time of two-pass classical Gram-Schmidt (``c = V^H w; w -= V c`` twice) against
    ``j`` columns, for every ``j`` in ``columns``: measured with NumPy BLAS at three column
    counts and integrated with the linear fit (the cost is affine in ``j``)."""

    from threadpoolctl import threadpool_limits

    if not columns:
        return 0.0
    top = max(columns)
    probe = sorted({max(1, top // 8), max(1, top // 2), max(1, top)})
    rng = np.random.default_rng(1)
    V = np.asarray(rng.normal(size=(top, rows)), dtype=dtype)
    if np.iscomplexobj(V):
        V = V + 1j * np.asarray(rng.normal(size=(top, rows)), dtype=np.float64)
    w0 = V[0].copy()
    times = []
    with threadpool_limits(limits=threads):
        for j in probe:
            Vj = V[:j]
            best = np.inf
            for _ in range(2):
                w = w0.copy()
                t0 = time.perf_counter()
                for _ in range(2):
                    c = (Vj @ w.conj()).conj()  # V^H w without a conjugated copy of V
                    w -= c @ Vj
                best = min(best, time.perf_counter() - t0)
            times.append(best)
    slope, intercept = np.polyfit(np.asarray(probe, dtype=float), np.asarray(times), 1) if len(probe) > 1 else (
        times[0] / probe[0], 0.0)
    return float(sum(max(intercept + slope * j, 0.0) for j in columns))


def _cpu_baseline(pipe, result) -> dict:
    """Estimated whole solve of the CPU restatement on this host: the NumPy port of the same
    plans timed on one residual, one JVP and one sigma per charge (single thread: the port's
    gathers do not thread), scaled by the GPU solve's operator counts, plus the Krylov
    orthogonalisation of the same solve (two-pass CGS in NumPy BLAS on every host thread the
    box grants, timed at three column counts and integrated over the solve's columns).  The
    small shifted solves are left out (optimistic for the CPU)."""

    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle.port import PortPlan

    sp = pipe.static
    p = sp.oracle
    t = result.amplitudes
    rng = np.random.default_rng(0)
    dt = p.tilde.unpack(rng.normal(size=p.num_variables))
    port = PortPlan(sp.lowered)
    t0 = time.perf_counter()
    port.load(0, t)
    port.apply(0, np.zeros(p.num_amplitudes, complex), sp.diagonal, t)
    t1 = time.perf_counter()
    port.load(1, dt)
    port.apply(2, np.zeros(p.num_amplitudes, complex), sp.diagonal, dt)
    t2 = time.perf_counter()
    sigma = {}
    for q, block in pipe.eom.items():
        pp = PortPlan(block.lowered)
        pp.load(0, t)
        x = rng.normal(size=block.dimension) + 1j * rng.normal(size=block.dimension)
        s0 = time.perf_counter()
        pp.load(2, x)
        pp.apply(3, np.zeros(block.dimension, complex), block.diagonal, x)
        sigma[q] = time.perf_counter() - s0
    t_res, t_jvp = t1 - t0, t2 - t1
    n_sigma = sum(sum(d) for d in result.krylov_dimensions.values())
    per_sigma = float(np.mean(list(sigma.values())))
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    restart = pipe.policy.gmres_restart
    gmres_columns = [j % restart + 1 for its in result.gmres_iterations for j in range(its)]
    arnoldi_columns = [j + 1 for dims in result.krylov_dimensions.values() for m in dims for j in range(m)]
    P = max(b.dimension for b in pipe.eom.values())
    t_gmres = _host_cgs_seconds(p.num_variables, gmres_columns, np.float64, threads)
    t_arnoldi = _host_cgs_seconds(P, arnoldi_columns, np.complex128, threads)
    solve_seconds = (result.residual_evaluations * t_res + result.jvp_evaluations * t_jvp + n_sigma * per_sigma
                     + t_gmres + t_arnoldi)
    return {
        "value": 1.0 / solve_seconds,
        "unit": "solves/sec",
        "cores": threads,
        "kind": "port",
        "sample": (
            f"NumPy port of the same plans on this host, single thread: 1 residual ({t_res:.2f}s), 1 JVP "
            f"({t_jvp:.2f}s), 1 sigma per charge ({per_sigma:.3f}s) at n={p.model.n_sites}, scaled by one GPU "
            f"solve's counts ({result.residual_evaluations} residuals, {result.jvp_evaluations} JVPs, {n_sigma} "
            f"sigmas); plus two-pass CGS of the same solve in NumPy BLAS on {threads} threads ({len(gmres_columns)} "
            f"GMRES steps: {t_gmres:.1f}s, {len(arnoldi_columns)} Arnoldi steps: {t_arnoldi:.1f}s, each timed at three "
            "column counts and integrated); shifted solves excluded (optimistic for the CPU)"
        ),
    }


def _jvp_probe(jvps: int, config: str = "siam24", precision: str = "fp64") -> int:
    """Counter workload: one residual of ``config`` (siam24 or kanamori39), ``jvps`` JVPs in
    ``precision``, then a 256 MiB copy for calibration."""

    import torch

    from c3lindblad_amd.configs import config2_model, config5_model
    from c3lindblad_amd.lowering import factorized_cc_problem
    from c3lindblad_amd.static import ComponentStaticProblem

    p = factorized_cc_problem(config5_model(0) if config == "kanamori39" else config2_model())
    dev = ComponentStaticProblem.compile(p).device
    if precision == "fp32":
        dev.plan.set_precision("D", "jvp", 32)
    rng = np.random.default_rng(0)
    x = torch.from_numpy(0.01 * rng.normal(size=p.num_variables)).cuda()
    r = torch.empty_like(x)
    dev.residual_real(x, r)
    dx = torch.from_numpy(rng.normal(size=p.num_variables)).cuda()
    for _ in range(jvps):
        dev.jvp_real(dx, r)
    a = torch.empty(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
    torch.empty_like(a).copy_(a)
    torch.cuda.synchronize()
    print("jvp probe done", flush=True)
    return 0


def _bench_batch(args) -> int:
    """The uneven-size batch: ``--instances`` instances of the 4..32-mode SIAM distribution (512 by
    default), each to ||pack R|| <= 1e-10 and both charges' G^R on 2048 frequencies, solved as
    groups of equal site count in lockstep (``c3lindblad_amd.batch``).  The instances go through in
    chunks of ``--batch-chunk``, in order of site count so that the groups fill: each chunk's host set-up (full HF, lowering, compilation, upload)
    and one untimed warm-up solve of its groups (device program images) precede its timed solve;
    ``value`` = instances solved / the sum of the timed regions."""

    import torch

    from c3lindblad_amd.batch import prepare_batch, solve_group
    from c3lindblad_amd.response import run_concurrently
    from c3lindblad_amd.configs import config3_grid, config3_model, config3_modes
    from c3lindblad_amd.solver import NewtonKrylovPolicy

    omega = config3_grid()
    policy = NewtonKrylovPolicy.mixed_precision(forcing_strategy="eisenstat_walker", gmres_restart=args.restart)
    # chunks of consecutive site counts, so that the groups inside a chunk are full
    ids = sorted(range(args.instances), key=lambda i: (config3_modes(i), i))
    timed, setup = 0.0, 0.0
    solutions, skipped, group_sizes, tuned = [], [], [], []
    steps = max(1, args.steps)
    for c0 in range(0, len(ids), args.batch_chunk):
        chunk = ids[c0:c0 + args.batch_chunk]
        prep = prepare_batch([config3_model(i) for i in chunk], indices=chunk, release_host=True,
                             workers=args.setup_workers,
                             max_group=args.max_group)
        setup += prep.setup_seconds
        skipped += prep.skipped
        group_sizes += [(g.n, g.count) for g in prep.groups]
        tuned += [r for g in prep.groups for r in g.tuning]
        for g in prep.groups:  # warm-up: program images are built at their first launch
            solve_group(g, omega, policy=policy)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            chunk_solutions = []
            if args.group_streams > 1:  # groups on concurrent streams, one host thread each
                tasks = [lambda g=g: solve_group(g, omega, policy=policy) for g in prep.groups]
                for sols in run_concurrently(tasks, streams=args.group_streams):
                    chunk_solutions += sols
            else:
                for g in prep.groups:
                    chunk_solutions += solve_group(g, omega, policy=policy)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        timed += dt
        solutions += chunk_solutions
        print(f"[bench] chunk {c0 // args.batch_chunk}: {len(chunk_solutions)} instances in {len(prep.groups)} groups, "
              f"setup {prep.setup_seconds:.1f}s, timed {dt:.2f}s for {steps} solve(s) of each", file=sys.stderr,
              flush=True)
        del prep
        torch.cuda.empty_cache()
    solved = len(solutions)
    n_of = {s.index: config3_modes(s.index) + 1 for s in solutions}
    line = {
        "metric": "CCSD+EOM G^R(ω) solves/sec, 1 MI355X: mixed+batched+learned vs FP64 @1e-8",
        "value": solved * steps / timed,
        "unit": "solves/sec",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 1,
        "ms_per_step": 1e3 * timed / steps,  # one step = one solve of every instance
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (JVP f32)",
        "data": "synthetic",
        "config": {
            "workload": f"siam_uneven_batch_{args.instances}_instances_4to32_modes_ccsd_eom_s13_gr_2048w",
            "variant": "mixed_fp32_jvp_grouped_lockstep",
            "jvp_matrix_cores": args.matrix_cores,
            "instances_requested": args.instances,
            "instances_solved": solved,
            "instances_skipped_no_full_hf": [i for i, _ in skipped],
            "groups": len(group_sizes),
            "group_sizes_by_n": sorted(group_sizes),
            "max_group": args.max_group,
            "groups_at_once": args.group_streams,
            "chunk": args.batch_chunk,
            "timed_seconds": timed,
            "autotuner": {p: {"launches_tuned": sum(1 for r in tuned if r["program"] == p),
                              "ms_default_sum": round(sum(r["ms_default"] for r in tuned if r["program"] == p), 3),
                              "ms_tuned_sum": round(sum(r["ms_tuned"] for r in tuned if r["program"] == p), 3)}
                          for p in ("residual", "jvp", "sigma")},
            "setup_seconds_total": setup,
            "n_sites_range": [min(n_of.values()), max(n_of.values())] if n_of else None,
            "omega_points": int(omega.size),
            "charges": [1, -1],
            "packed_residual_max": max(s.packed_residual_norm for s in solutions),
            "max_shift_relative_residual": max(s.max_shift_residual for s in solutions),
            "jvp_evaluations_total": int(sum(s.jvp_evaluations for s in solutions)),
            "newton_iterations_max": int(max(s.newton_iterations for s in solutions)),
            "layers_in_timed_region": {"mixed_precision": True,
                                       "batched": "instances of equal site count in lockstep groups (union plans); "
                                                  "every frequency on one Krylov basis per channel",
                                       "learned_warm_start": False},
        },
    }
    print(json.dumps(line), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="siam24", choices=("siam24", "readme", "kanamori39", "batch512", "siam12"))
    ap.add_argument("--space", default="S13", choices=("S13", "S135"),
                    help="EOM response space: S13, or S135 (S1 + S3 + the rank-five shell)")
    ap.add_argument("--instances", type=int, default=512, help="batch512: instances of the uneven-size batch")
    ap.add_argument("--batch-chunk", type=int, default=64, help="batch512: instances set up and solved per chunk")
    ap.add_argument("--max-group", type=int, default=2, help="batch512: instances per lockstep group")
    ap.add_argument("--group-streams", type=int, default=4,
                    help="batch512: groups solved at once, each on its own stream and host thread")
    ap.add_argument("--setup-workers", type=int, default=min(16, os.cpu_count() or 1),
                    help="batch512: host processes for the instances' set-up (outside the timed region)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--forcing", default="eisenstat_walker", choices=("legacy", "eisenstat_walker"),
                    help="inexact-Newton forcing strategy of the CCSD solve (the reference's two options)")
    ap.add_argument("--restart", type=int, default=150, help="GMRES restart length of the CCSD solve")
    ap.add_argument("--batch", type=int, default=1,
                    help="independent instances of the workload's distribution solved concurrently (seeds 0..B-1)")
    ap.add_argument("--no-plain-estimate", action="store_true",
                    help="skip the sampled estimate of the per-frequency FP64 path (needs the FP64 reference solve)")
    ap.add_argument("--no-fp64-reference", action="store_true",
                    help="skip the extra solve with the reference's default FP64 policy")
    ap.add_argument("--precision", default="fp32", choices=("fp64", "fp32"),
                    help="precision of the Jacobian-vector products inside GMRES (fp32 = mixed precision)")
    ap.add_argument("--eom-streams", type=int, default=None,
                    help="streams for the EOM resolvents (default: one per resolvent)")
    ap.add_argument("--serial-probe", action="store_true",
                    help="also time one solve with the EOM channels on one stream (no instrumentation)")
    ap.add_argument("--calibration-copy", action="store_true",
                    help="end with one 256 MiB device copy (known bytes for hardware-counter calibration)")
    ap.add_argument("--matrix-cores", default="auto", choices=("auto", "on", "off"),
                    help="FP32 JVPs: rank-4 two-atom blocks on the matrix cores where they gain (auto), "
                         "always (on) or never (off)")
    ap.add_argument("--jvp-probe", type=int, default=0, metavar="K",
                    help="profiling workload only: one residual and K JVPs (--config siam24 or kanamori39, "
                         "--precision), then a 256 MiB copy")
    args = ap.parse_args()
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from c3lindblad_amd.device import DevicePlan

    DevicePlan.matrix_cores = {"off": 0, "auto": 1, "on": 2}[args.matrix_cores]
    if args.jvp_probe:
        return _jvp_probe(args.jvp_probe, args.config, args.precision)
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(
            "bench.py runs the single-GPU path only: this build has no multi-GPU mode "
            "(the solve is one instance on one MI355X); run with --gpus 1",
            file=sys.stderr,
        )
        return 2
    if args.config == "batch512":
        return _bench_batch(args)
    import torch

    from c3lindblad_amd.configs import config2_grid, config2_model, readme_grid, readme_model
    from c3lindblad_amd.device import KernelTimer
    from c3lindblad_amd.pipeline import SolvePipeline, solve_concurrently
    from c3lindblad_amd.plan_arrays import PROGRAMS

    if args.config == "siam24":
        model, grid, charges = config2_model(), config2_grid(), (1, -1)
        workload = "siam_24_pseudomodes_U4_ccsd_eom_s13_gr_4001w"
    elif args.config == "siam12":
        from c3lindblad_amd.configs import siam_model

        model, grid, charges = siam_model(11, 4.0, seed=0), config2_grid(), (1, -1)
        workload = "siam_11_pseudomodes_U4_ccsd_eom_" + args.space.lower() + "_gr_4001w"
    elif args.config == "kanamori39":
        from c3lindblad_amd.configs import config5_grid, config5_model

        model, grid, charges = config5_model(0), config5_grid(), (1, -1)
        workload = "kanamori_3orb_U3_J0.5_Vpqrs_36_pseudomodes_ccsd_eom_s13_gr_8001w"
    else:
        model, grid, charges = readme_model(), readme_grid(), (1,)
        workload = "readme_3site_U1.2_ccsd_eom_s13_gr_81w"
    from c3lindblad_amd.solver import NewtonKrylovPolicy

    if args.precision == "fp32":
        policy = NewtonKrylovPolicy.mixed_precision(forcing_strategy=args.forcing, gmres_restart=args.restart)
    else:
        policy = NewtonKrylovPolicy(forcing_strategy=args.forcing, gmres_restart=args.restart)
    pipe = SolvePipeline.prepare(model, grid, 0.0, charges=charges, policy=policy, space=args.space)
    if args.space == "S135" and args.config == "siam24":
        workload = "siam_24_pseudomodes_U4_ccsd_eom_s135_gr_4001w"
    pipe.eom_streams = args.eom_streams
    print(f"[bench] setup {pipe.setup_seconds:.2f}s n={model.n_sites} D={pipe.static.oracle.num_amplitudes}", file=sys.stderr, flush=True)
    extra = []
    if args.batch > 1:
        if args.config != "siam24":
            print("--batch applies to the siam24 distribution only", file=sys.stderr)
            return 2
        seed = 1
        while len(extra) < args.batch - 1:
            try:  # instances whose full-HF reference does not converge are rejected, as by the reference
                extra.append(SolvePipeline.prepare(config2_model(seed), grid, 0.0, charges=charges, policy=policy))
            except RuntimeError as err:
                print(f"[bench] instance {seed} skipped: {err}", file=sys.stderr, flush=True)
            seed += 1
    pipes = [pipe] + extra
    for k in range(args.warmup):
        if len(pipes) > 1:
            solve_concurrently(pipes, 1)
        r = pipe.solve()
        print(f"[bench] warmup {k} {r.seconds}", file=sys.stderr, flush=True)
    plans = [pipe.static.device.plan] + [b.device_plan() for b in pipe.eom.values()]
    batch_results = None
    if len(pipes) > 1:
        # timed region: every instance solved `steps` times, all instances concurrently
        batch_results, elapsed = solve_concurrently(pipes, args.steps)
        results = batch_results[0]
        print(f"[bench] batch of {len(pipes)}: {elapsed:.3f}s for {len(pipes) * args.steps} solves", file=sys.stderr, flush=True)
        timed_solves = len(pipes) * args.steps
    else:
        torch.cuda.synchronize()
        start = time.perf_counter()
        results = []
        for k in range(args.steps):
            results.append(pipe.solve())
            print(f"[bench] step {k} {results[-1].seconds}", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - start
        timed_solves = args.steps
    # Kernel timing: one extra solve with every launch on one stream (the lockstep EOM
    # basis pairs of both charge blocks interleaved on it, as in the timed solves but not
    # overlapped), every plan program and every driver orthogonalisation call bracketed by
    # HIP events (events of concurrent streams would also count the other streams' work).
    serial_phases = None
    if args.serial_probe:
        pipe.concurrent = False
        serial_phases = pipe.solve().seconds
        pipe.concurrent = True
        print(f"[bench] serial EOM solve {serial_phases}", file=sys.stderr, flush=True)
    lib = _lib_handle()
    for plan in plans:
        plan.set_timing(True)
    lib.c3l_krylov_set_timing(1)
    KernelTimer.enabled = True
    pipe.eom_streams = 1
    torch.cuda.synchronize()
    instrumented = time.perf_counter()
    pipe.solve()
    torch.cuda.synchronize()
    instrumented = time.perf_counter() - instrumented
    pipe.eom_streams = args.eom_streams
    KernelTimer.enabled = False
    kernels = {name: d for name, d in KernelTimer.summary().items()}
    for plan in plans:
        for name, pid in PROGRAMS.items():
            ms, launches = plan.program_time(name)
            if launches == 0:
                continue
            d = kernels.setdefault(f"wedge_program[{name}]", {"ms": 0.0, "launches": 0, "bytes": 0.0})
            d["ms"] += ms
            d["launches"] += launches
            atom_bytes = 8 if (name == "jvp" and args.precision == "fp32") else 16
            d["bytes"] += launches * _algorithmic_program_bytes(plan.lowered, pid, atom_bytes)
        plan.set_timing(False)
    import ctypes

    for family, name in enumerate(("reduce_kernel<double2>[arnoldi]", "update_kernel<double2>[arnoldi]",
                                   "reduce_kernel<double>[gmres]", "update_kernel<double>[gmres]")):
        ms, launches, nbytes = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double()
        lib.c3l_krylov_timing(family, ctypes.byref(ms), ctypes.byref(launches), ctypes.byref(nbytes))
        if launches.value:
            kernels[name] = {"ms": ms.value, "launches": launches.value, "bytes": nbytes.value}
    lib.c3l_krylov_set_timing(0)
    dominant = max(kernels, key=lambda k: kernels[k]["ms"])
    traffic, traffic_source = _measured_traffic(dominant)
    dk = kernels[dominant]
    achieved = dk["bytes"] / (dk["ms"] * 1e-3) / 1e9
    last = results[-1]
    spread = max(
        float(np.max(np.abs(r.green[q] - last.green[q])) / np.max(np.abs(last.green[q]))) for r in results for q in last.green
    )
    line = {
        "metric": "CCSD+EOM G^R(ω) solves/sec, 1 MI355X: mixed+batched+learned vs FP64 @1e-8",
        "value": timed_solves / elapsed,
        "unit": "solves/sec",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,  # one step = one solve of every instance of the batch
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64 (JVP f32)",
        "data": "synthetic",
        "config": {
            "workload": workload,
            "variant": "fp64_gpu" if args.precision == "fp64" else "mixed_fp32_jvp",
            "jvp_matrix_cores": args.matrix_cores if args.precision == "fp32" else "off",
            "layers_in_timed_region": {
                "mixed_precision": args.precision == "fp32",
                "batched": "every frequency on one Krylov basis per channel; addition and removal bases of a "
                           "charge block in lockstep; charge blocks on concurrent streams"
                           + (f"; {len(pipes)} instances on concurrent streams" if len(pipes) > 1 else ""),
                "learned_warm_start": False,
            },
            "newton_policy": {"forcing": args.forcing, "gmres_restart": args.restart,
                              "jacobian_precision": args.precision,
                              "minimum_gmres_relative_tolerance": policy.minimum_gmres_relative_tolerance,
                              "maximum_gmres_relative_tolerance": policy.maximum_gmres_relative_tolerance,
                              "initial_gmres_relative_tolerance": policy.initial_gmres_relative_tolerance,
                              "eisenstat_walker_gamma": policy.eisenstat_walker_gamma,
                              "eisenstat_walker_exponent": policy.eisenstat_walker_exponent},
            "n_sites": model.n_sites,
            "ccsd_amplitudes": pipe.static.oracle.num_amplitudes,
            "eom_dimension_per_charge": {str(q): b.dimension for q, b in pipe.eom.items()},
            "omega_points": int(grid.size),
            "eta": 0.0,
            "charges": list(charges),
            "setup_seconds": pipe.setup_seconds,
            "autotuner": [r if r["program"] == "residual_path" else
                          {"program": r["program"] + (f"[{r['charge']:+d}]" if "charge" in r else "")
                           + (f"[fp{r['bits']}]" if "bits" in r else ""),
                           "us_default": round(1e3 * r["ms_default"], 2), "us_tuned": round(1e3 * r["ms_tuned"], 2),
                           "shape": r["shape"]} for r in pipe.tuning],
            "newton_iterations": last.newton_iterations,
            "jvp_evaluations": last.jvp_evaluations,
            "residual_evaluations": last.residual_evaluations,
            "krylov_dimensions": {str(q): list(v) for q, v in last.krylov_dimensions.items()},
            "packed_residual_norm": last.packed_residual_norm,
            "max_shift_relative_residual": last.max_shift_residual,
            "gr_max_rel_spread_between_steps": spread,
            "seconds_per_phase": last.seconds,
            **({"serial_eom_seconds_per_phase": serial_phases} if serial_phases else {}),
            "batch_instances": len(pipes),
            "batch_packed_residual_max": max(r.packed_residual_norm for rs in (batch_results or [results]) for r in rs),
            "batch_max_shift_residual": max(r.max_shift_residual for rs in (batch_results or [results]) for r in rs),
        },
        "roofline": {
            "kernel": dominant,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": (f"committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary of the same kernel and workload "
                               f"({traffic_source}), not measured in this run") if traffic_source else None,
            "avg_launch_us": 1e3 * dk["ms"] / dk["launches"],
            "launches": dk["launches"],
            "share_of_step": dk["ms"] * 1e-3 / instrumented,
            "measured_on": "one extra instrumented solve after the timed steps (all launches on one stream)",
        },
        "kernels": {
            k: {"ms": round(v["ms"], 3), "launches": v["launches"],
                "algorithmic_GBps": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1) if v["ms"] > 0 else None}
            for k, v in sorted(kernels.items(), key=lambda kv: -kv[1]["ms"])[:8]
        },
    }
    if not args.no_fp64_reference:
        import dataclasses

        ref_pipe = dataclasses.replace(pipe, policy=NewtonKrylovPolicy())
        ref_pipe.solve()
        torch.cuda.synchronize()
        t_ref = time.perf_counter()
        ref = ref_pipe.solve()
        torch.cuda.synchronize()
        t_ref = time.perf_counter() - t_ref
        line["config"]["fp64_reference_policy"] = {
            "policy": "FP64 JVPs, legacy forcing, GMRES restart 40 (reference defaults), same shared-basis resolvent",
            "solves_per_sec": 1.0 / t_ref,
            "speedup_of_this_variant": line["value"] * t_ref,
            "jvp_evaluations": ref.jvp_evaluations,
            "packed_residual_norm": ref.packed_residual_norm,
            "gr_rel_error_of_this_variant": max(
                float(np.max(np.abs(last.green[q] - ref.green[q])) / np.max(np.abs(ref.green[q]))) for q in ref.green
            ),
        }
        if not args.no_plain_estimate:
            plain = _plain_fp64_estimate(ref_pipe, ref)
            plain["speedup_of_this_variant"] = line["value"] / plain["solves_per_sec"]
            line["config"]["plain_fp64_gpu_path"] = plain
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = _cpu_baseline(pipe, last)
    if args.calibration_copy:
        a = torch.empty(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
        torch.empty_like(a).copy_(a)
        torch.cuda.synchronize()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

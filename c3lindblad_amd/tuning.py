"""Set-up step of a solve: the launch-shape autotuner on every program a solve launches.

The reference leaves its schedules to XLA (``cc/component_jax.py:248-462``).  Here each
program image (the coordinate tiles of ``wedge_program_kernel`` and, for FP32 programs, the
matrix-core tiles of ``tile_program_mfma_kernel``) has a launch shape: lanes per coordinate,
coordinates per thread, the workgroup target behind both, the XCD slicing of a block's tiles
and the matrix-core entry threshold.  ``c3l_plan_autotune`` times candidates on the plan itself
(coordinate descent from the default heuristics) and keeps the fastest for every plan with the
same digest, so twin plans and repeated instances share one tuning.  Tuning runs once per plan
at set-up, outside any timed solve; programs are timed on zero inputs (their cost does not
depend on the values).  The set-up also picks, per static problem, how the residual is
evaluated: by its program, by Gauss-Legendre quadrature of FP64 JVPs or by Simpson's rule
(two FP64 JVPs and the low-degree program), whichever is fastest.
"""

from __future__ import annotations

import numpy as np

from . import _lib


def autotune_static(static, jvp_bits: int = 64, reps: int = 5) -> list[dict]:
    """Tune the residual program and the JVP program (in the precision the Newton solve uses)
    of a compiled static problem, then choose how the residual is evaluated
    (``choose_residual_path``)."""

    dev = static.device
    dev.t.zero_()
    dev.dt.zero_()
    dev.load_amplitudes(dev.t)
    out = [dev.plan.autotune("residual", dev.out, dev.diag, dev.t, reps=reps)]
    dev.plan.load("D", dev.dt)
    out.append(dev.plan.autotune("jvp", dev.out, dev.diag, dev.dt, reps=reps))  # FP64: the quadrature path
    if jvp_bits == 32:
        dev.plan.set_precision("D", "jvp", 32)
        try:
            dev.plan.load("D", dev.dt)
            out.append(dict(dev.plan.autotune("jvp", dev.out, dev.diag, dev.dt, reps=reps), bits=32))
        finally:
            dev.plan.set_precision("D", "jvp", 64)
    out.append(choose_residual_path(static))
    return out


def choose_residual_path(static, reps: int = 3) -> dict:
    """Time one residual evaluation (``StaticDevice.residual_at``: the amplitude loads with their
    creation tensors that the path needs, then the residual) by the residual program, by Gauss-Legendre quadrature of FP64 JVPs
    (``StaticDevice._residual_by_quadrature``) and, for residuals of degree at most four, by
    Simpson's rule (``StaticDevice._residual_by_simpson``), and keep the fastest."""

    torch = _lib.require_device()
    dev = static.device
    rng = np.random.default_rng(0)
    t = 0.01 * (rng.normal(size=dev.dim) + 1j * rng.normal(size=dev.dim))
    dev.t.copy_(torch.from_numpy(t).to(_lib.DEVICE))
    times = {}
    paths = ("program", "quadrature", "simpson") if static.residual_degree <= 4 else ("program", "quadrature")
    for path in paths:
        dev.residual_path = path
        dev.residual_at(dev.t, dev.out)  # untimed: first use builds the constant part R(0)
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        for _ in range(reps):
            dev.residual_at(dev.t, dev.out)
        stop.record()
        stop.synchronize()
        times[path] = start.elapsed_time(stop) / reps
    dev.residual_path = min(times, key=times.get)
    return {"program": "residual_path", "ms_program": times["program"], "ms_quadrature": times["quadrature"],
            "ms_simpson": times.get("simpson"), "choice": dev.residual_path, "quadrature_nodes": int(len(dev.quadrature_nodes()[0])),
            "residual_degree": int(static.residual_degree)}


def autotune_sigma(block, reps: int = 5) -> dict:
    """Tune the sigma program of a compiled EOM charge block (shared by its twin slots)."""

    torch = _lib.require_device()
    plan = block.device_plan(0)
    D = block.static.oracle.num_amplitudes
    P = block.dimension
    t = torch.zeros(D, dtype=torch.complex128, device=_lib.DEVICE)
    x = torch.zeros(P, dtype=torch.complex128, device=_lib.DEVICE)
    out = torch.empty_like(x)
    diag = torch.from_numpy(block.diagonal).to(_lib.DEVICE)
    plan.load("T", t)
    plan.load("X", x)
    return plan.autotune("sigma", out, diag, x, reps=reps)

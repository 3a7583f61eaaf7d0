"""Groups of model instances solved in lockstep (``c3lindblad_amd.batch``) against the same
instances solved one at a time by the single-instance path, which the other GPU tests pin
against the CPU oracle.

The group mixes an interacting and a U = 0 instance, so its union plan joins plans of different
structure.  Bars: every instance's packed residual <= 1e-10 (checked again with the NumPy port),
the same Newton iteration and JVP counts as its single solve, amplitudes within 1e-10 and G^R
within 1e-10 relative of the single solve, every shift residual <= 1e-10.
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _models():
    from c3lindblad_amd.configs import siam_model

    return [siam_model(7, 3.0, seed=11), siam_model(7, 0.0, seed=12), siam_model(7, 5.0, seed=13, epsilon_d=-2.0)]


def test_group_solve_equals_single_instance_solves(torch_gpu):
    from c3lindblad_amd.batch import InstanceGroup, compile_instance, solve_group
    from c3lindblad_amd.eom import ComponentSigmaOperator, component_eom_context, component_eom_sources
    from c3lindblad_amd.response import retarded_green_blocks
    from c3lindblad_amd.solver import NewtonKrylovPolicy, solve_component_ccsd
    from oracle.port import PortPlan

    print("case group of three n=8 instances, 2048 frequencies", flush=True)
    omega = np.linspace(-8.0, 8.0, 2048)
    policy = NewtonKrylovPolicy.mixed_precision(forcing_strategy="eisenstat_walker", gmres_restart=150)
    members = [compile_instance(i, m) for i, m in enumerate(_models())]
    group = InstanceGroup(members, (1, -1))
    group.upload()
    solved = solve_group(group, omega, policy=policy)
    for inst, got in zip(members, solved):
        sp = inst.static
        single = solve_component_ccsd(sp, policy=policy)
        print(f"  instance {inst.index}: group newton={got.newton_iterations} jvps={got.jvp_evaluations} "
              f"packed={got.packed_residual_norm:.2e}; single newton={single.iterations} jvps={single.jvp_evaluations}",
              flush=True)
        assert got.converged and got.packed_residual_norm <= 1e-10
        assert got.jvp_evaluations == single.jvp_evaluations
        assert _rel(got.amplitudes, single.amplitudes) < 1e-10
        port = PortPlan(sp.lowered)
        port.load(0, got.amplitudes, creation=False)
        R = port.apply(0, np.zeros(sp.oracle.num_amplitudes, complex), sp.diagonal, got.amplitudes)
        assert np.linalg.norm(sp.oracle.tilde.pack(R, check=False)) <= 1e-10
        pairs = []
        for q in (1, -1):
            ctx = component_eom_context(inst.eom[q], single.amplitudes)
            pairs.append((ComponentSigmaOperator(ctx), component_eom_sources(ctx)))
        for q, ref in zip((1, -1), retarded_green_blocks(pairs, omega, 0.0, rtol=1e-10)):
            assert _rel(got.green[q], ref.retarded) < 1e-10, q
        assert got.max_shift_residual <= 1e-10


def test_group_gmres_equals_single_gmres(torch_gpu):
    """One Newton correction: c3l_gmres_group on three systems against c3l_gmres on each (FP64
    basis and JVPs): equal iteration counts and steps within 1e-12."""

    import ctypes

    import torch

    from c3lindblad_amd import _lib
    from c3lindblad_amd.batch import InstanceGroup, compile_instance
    from c3lindblad_amd.solver import _DeviceNewton

    print("case group gmres", flush=True)
    members = [compile_instance(i, m) for i, m in enumerate(_models())]
    group = InstanceGroup(members, (1,))
    group.upload()
    G, D = group.count, group.D
    rng = np.random.default_rng(2)
    xs = [0.01 * rng.normal(size=D) for _ in members]
    bs = [rng.normal(size=D) for _ in members]
    singles = []
    for inst, x0, b0 in zip(members, xs, bs):
        state = _DeviceNewton(inst.static, 40, 64)
        x = torch.from_numpy(x0).cuda()
        state.residual(x, state.r)  # linearisation point
        b = torch.from_numpy(b0).cuda()
        step = torch.empty_like(b)
        info, its, true = state.gmres(b, step, 1e-8, 0.0, 40, 6)
        singles.append((info, its, true, step.cpu().numpy()))
    plan, tilde = group._dev["plan"], group._dev["tilde"]
    dev = "cuda"
    t = torch.zeros(G * D, dtype=torch.complex128, device=dev)
    tilde.unpack(torch.from_numpy(np.concatenate(xs)).cuda(), t)
    plan.load("T", t)
    b = torch.from_numpy(np.concatenate(bs)).cuda()
    step = torch.empty_like(b)
    w, tmp = torch.empty_like(b), torch.empty_like(b)
    dt = torch.zeros(G * D, dtype=torch.complex128, device=dev)
    out = torch.zeros_like(dt)
    V = torch.zeros(41, G * D, dtype=torch.float64, device=dev)
    scratch_size, coef_size = _lib.krylov_workspace(False, D, 40)
    scratch = torch.zeros(G * scratch_size, dtype=torch.float64, device=dev)
    coef = torch.zeros(G * coef_size, dtype=torch.float64, device=dev)
    maps = _lib.TildeMaps(tilde.dim, tilde.real_dim, _lib.dptr(tilde.k0), _lib.dptr(tilde.k1), _lib.dptr(tilde.ab),
                          _lib.dptr(tilde.src), _lib.dptr(tilde.w), _lib.dptr(group._dev["inv_diag"]))
    rt = (ctypes.c_double * G)(*([1e-8] * G))
    at = (ctypes.c_double * G)(*([0.0] * G))
    active = (ctypes.c_int32 * G)(*([1] * G))
    info, its, calls = (ctypes.c_int32 * G)(), (ctypes.c_int32 * G)(), (ctypes.c_int32 * G)()
    true = (ctypes.c_double * G)()
    _lib.call("c3l_gmres_group", plan.handle, ctypes.byref(maps), G, _lib.dptr(group._dev["diag"]), _lib.dptr(dt),
              _lib.dptr(out), _lib.dptr(b), _lib.dptr(step), _lib.dptr(V), 64, _lib.dptr(w), _lib.dptr(tmp),
              _lib.dptr(scratch), _lib.dptr(coef), rt, at, active, 40, 6, info, its, true, calls,
              _lib.stream_ptr(torch))
    got = step.cpu().numpy()
    for i, (s_info, s_its, s_true, s_step) in enumerate(singles):
        print(f"  system {i}: group its={its[i]} info={info[i]} true={true[i]:.2e}; single its={s_its} "
              f"info={s_info} true={s_true:.2e}", flush=True)
        assert (info[i], its[i]) == (s_info, s_its)
        assert abs(true[i] - s_true) <= 1e-12
        assert _rel(got[i * D:(i + 1) * D], s_step) < 1e-12


def test_groups_on_concurrent_streams_equal_sequential_solves(torch_gpu):
    """Four two-instance groups solved at once on four streams, one host thread each
    (``run_concurrently``, as ``bench.py --config batch512`` does), against the same groups solved
    one after the other: every instance's Newton and JVP counts equal, amplitudes and G^R equal
    bit for bit (the streams share nothing but the device)."""

    from c3lindblad_amd.batch import InstanceGroup, compile_instance, solve_group
    from c3lindblad_amd.configs import siam_model
    from c3lindblad_amd.response import run_concurrently
    from c3lindblad_amd.solver import NewtonKrylovPolicy

    print("case four groups of two n=6 instances on concurrent streams, 512 frequencies", flush=True)
    omega = np.linspace(-8.0, 8.0, 512)
    policy = NewtonKrylovPolicy.mixed_precision(forcing_strategy="eisenstat_walker", gmres_restart=150)
    models = [siam_model(5, 1.0 + 0.5 * k, seed=40 + k) for k in range(8)]
    members = [compile_instance(i, m) for i, m in enumerate(models)]
    groups = [InstanceGroup(members[2 * g:2 * g + 2], (1, -1)) for g in range(4)]
    for g in groups:
        g.upload()
    sequential = [solve_group(g, omega, policy=policy) for g in groups]
    tasks = [lambda g=g: solve_group(g, omega, policy=policy) for g in groups]
    concurrent = run_concurrently(tasks, streams=4)
    for seq, con in zip(sequential, concurrent):
        for a, b in zip(seq, con):
            assert a.converged and b.converged
            assert (a.newton_iterations, a.jvp_evaluations) == (b.newton_iterations, b.jvp_evaluations)
            assert np.array_equal(a.amplitudes, b.amplitudes)
            for q in (1, -1):
                assert np.array_equal(a.green[q], b.green[q]), q

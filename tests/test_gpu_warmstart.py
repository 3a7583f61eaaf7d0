"""Learned warm start on the GPU: kernels against the NumPy oracle (oracle/mlp.py)
and the gain of a trained network over the zero initial guess.

Bars: bf16 GEMM (FP32 accumulation) within 1e-5 relative of the float64 product of
the same bf16 values; MLP forward/gradients within 2e-2 relative (max-norm) of the
bf16-emulating oracle (single bf16 rounding flips of near-tie activations are allowed);
Adam within 1e-6 relative.
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


@pytest.mark.parametrize("M,N,K,layout", [(70, 45, 37, "nn"), (64, 128, 48, "tn"), (33, 17, 200, "nt")])
def test_bf16_gemm_matches_float64_product(torch_gpu, M, N, K, layout):
    from c3lindblad_amd import _lib
    from oracle import mlp

    print(f"case bf16 gemm {M}x{N}x{K} {layout}", flush=True)
    torch = torch_gpu
    rng = np.random.default_rng(M + N + K)
    A = mlp.bf16(rng.normal(size=(M, K)).astype(np.float32))
    B = mlp.bf16(rng.normal(size=(K, N)).astype(np.float32))
    # stored layouts: "t" keeps the transpose in memory and reads it through strides
    As = A.T.copy() if layout[0] == "t" else A
    Bs = B.T.copy() if layout[1] == "t" else B
    a_rs, a_cs = (1, M) if layout[0] == "t" else (K, 1)
    b_rs, b_cs = (1, K) if layout[1] == "t" else (N, 1)
    Ad = torch.from_numpy(mlp.bits(As).view(np.int16)).cuda()
    Bd = torch.from_numpy(mlp.bits(Bs).view(np.int16)).cuda()
    C0 = rng.normal(size=(M, N)).astype(np.float32)
    Cd = torch.from_numpy(C0.copy()).cuda()
    _lib.call("c3l_bf16_gemm", _lib.dptr(Ad), a_rs, a_cs, _lib.dptr(Bd), b_rs, b_cs, _lib.dptr(Cd), N, M, N, K, 1.5, 0.5,
              _lib.stream_ptr(torch))
    want = 1.5 * (A.astype(np.float64) @ B.astype(np.float64)) + 0.5 * C0
    assert _rel(Cd.cpu().numpy(), want) < 1e-5


def test_mlp_forward_backward_and_adam_match_oracle(torch_gpu):
    from c3lindblad_amd.warmstart import AmplitudeMLP
    from oracle import mlp

    print("case mlp forward/backward/adam", flush=True)
    torch = torch_gpu
    rng = np.random.default_rng(9)
    sizes = (40, 64, 64, 64, 8)
    batch = 16
    W = [(rng.normal(size=(k, n)) / np.sqrt(k)).astype(np.float32) for k, n in zip(sizes[:-1], sizes[1:])]
    b = [(0.1 * rng.normal(size=n)).astype(np.float32) for n in sizes[1:]]
    x = rng.normal(size=(batch, sizes[0])).astype(np.float32)
    t = rng.normal(size=(batch, sizes[-1])).astype(np.float32)
    net = AmplitudeMLP(sizes)
    net.set_weights(W, b)
    x16 = net.to_bf16(torch.from_numpy(x).cuda())
    y, cache = net.forward(x16, batch)
    want_y, _ = mlp.forward(W, b, x)
    assert _rel(y.cpu().numpy(), want_y) < 2e-2
    loss, dy, dy16 = net.loss_gradient(y, torch.from_numpy(t).cuda(), batch)
    gW, gb = net.backward(cache, dy, dy16, batch)
    want_loss, want_gW, want_gb = mlp.loss_and_backward(W, b, x, t)
    assert abs(float(loss.sum().item()) / (batch * sizes[-1]) - want_loss) < 2e-2 * want_loss
    for l in range(len(W)):
        assert _rel(gW[l].cpu().numpy(), want_gW[l]) < 2e-2
        assert _rel(gb[l].cpu().numpy(), want_gb[l]) < 2e-2
    # one Adam step on the device gradients against the NumPy update of the same gradients
    g0 = [g.cpu().numpy() for g in gW]
    net.adam(gW, gb, 1e-3)
    for l in range(len(W)):
        w1, _, _ = mlp.adam(W[l].astype(np.float64), 0.0, 0.0, g0[l].astype(np.float64), 1, 1e-3)
        assert _rel(net.W[l].cpu().numpy(), w1) < 1e-6


def test_learned_warm_start_against_zero_guess(torch_gpu):
    """Train on converged solutions of small SIAMs drawn from one distribution, solve
    held-out instances from the network's initial amplitudes, and report the initial
    residual and JVP counts against the zero guess.  Newton must reach the same 1e-10
    root; the held-out blend keeps the guess from doing clearly worse than zero."""

    from c3lindblad_amd.configs import siam_model
    from c3lindblad_amd.lowering import factorized_cc_problem
    from c3lindblad_amd.solver import solve_component_ccsd
    from c3lindblad_amd.static import ComponentStaticProblem
    from c3lindblad_amd.warmstart import WarmStart

    print("case learned warm start (n=6 SIAMs)", flush=True)
    rng = np.random.default_rng(4)
    models, problems = [], []
    seed = 7000
    while len(models) < 48:
        m = siam_model(5, float(rng.uniform(1.0, 3.0)), seed=seed)
        seed += 1
        try:  # instances without a converged full-HF reference are rejected, as by the reference
            problems.append(ComponentStaticProblem.compile(factorized_cc_problem(m)))
        except RuntimeError:
            continue
        models.append(m)
    sols = [solve_component_ccsd(p) for p in problems]
    assert all(s.converged for s in sols)
    train, test = list(range(40)), list(range(40, 48))
    ws = WarmStart.fit([problems[i].oracle for i in train], [sols[i].amplitudes for i in train], rank=16, epochs=300,
                       batch=16)
    assert ws.train_losses[-1] < 0.5 * ws.train_losses[0]
    assert 0.0 <= ws.blend <= 1.0
    zero_r, warm_r, zero_j, warm_j = [], [], [], []
    for i in test:
        warm = solve_component_ccsd(problems[i], initial_variables=ws.initial_variables(problems[i].oracle))
        assert warm.converged and warm.packed_residual_norm <= 1e-10
        assert np.max(np.abs(warm.amplitudes - sols[i].amplitudes)) < 1e-8 * max(1.0, np.max(np.abs(sols[i].amplitudes)))
        zero_r.append(sols[i].history[0].residual_norm)
        warm_r.append(warm.history[0].residual_norm if warm.history else warm.packed_residual_norm)
        zero_j.append(sols[i].jvp_evaluations)
        warm_j.append(warm.jvp_evaluations)
    print(f"blend {ws.blend:.3f}; initial residual: zero guess {np.mean(zero_r):.3e}, warm start {np.mean(warm_r):.3e}; "
          f"JVPs: zero guess {np.mean(zero_j):.1f}, warm start {np.mean(warm_j):.1f}", flush=True)
    assert np.mean(warm_r) <= 1.5 * np.mean(zero_r)


def test_warm_start_gain_ceiling_on_the_siam24_distribution(torch_gpu):
    """What an initial guess must achieve before it saves Newton-Krylov work, on the 24-mode SIAM
    (n = 25, the workload the learned warm start is meant for): Newton from the converged root
    moved by a relative distance eps (a random direction, and a shrink toward zero as the
    network's held-out blend produces), mixed precision with Eisenstat-Walker forcing.  Every
    solve reaches the same root (packed residual <= 1e-10, amplitudes within 1e-8), and no guess
    1 % or more from the root saves 10 % of the JVPs of the zero guess.  The printed table is the
    gain ceiling of any predictor of a given accuracy."""

    from c3lindblad_amd.configs import config2_model
    from c3lindblad_amd.lowering import factorized_cc_problem
    from c3lindblad_amd.solver import NewtonKrylovPolicy, solve_component_ccsd
    from c3lindblad_amd.static import ComponentStaticProblem

    print("case warm-start ceiling siam24", flush=True)
    p = factorized_cc_problem(config2_model(0))
    sp = ComponentStaticProblem.compile(p)
    policy = NewtonKrylovPolicy.mixed_precision(forcing_strategy="eisenstat_walker", gmres_restart=150)
    zero = solve_component_ccsd(sp, policy=policy)
    assert zero.converged
    xs = p.tilde.pack(zero.amplitudes, check=False)
    scale = max(1.0, float(np.max(np.abs(zero.amplitudes))))
    rng = np.random.default_rng(5)
    u = rng.normal(size=xs.size)
    u *= np.linalg.norm(xs) / np.linalg.norm(u)
    rows = [("zero guess", 1.0, zero)]
    for eps in (0.3, 0.1, 0.03, 0.01, 1e-3):
        for kind, x0 in (("random", xs + eps * u), ("shrink", (1.0 - eps) * xs)):
            sol = solve_component_ccsd(sp, policy=policy, initial_variables=x0)
            assert sol.converged and sol.packed_residual_norm <= 1e-10
            assert np.max(np.abs(sol.amplitudes - zero.amplitudes)) < 1e-8 * scale
            rows.append((kind, eps, sol))
    for kind, eps, sol in rows:
        print(f"  {kind:10s} eps={eps:<6g} newton={sol.iterations} jvps={sol.jvp_evaluations} "
              f"gmres={[h.gmres_iterations for h in sol.history]} fallback={sol.warm_start_fallback}", flush=True)
    # the measured ceiling: a guess 1 % or more from the root (the learned network's held-out
    # error is far larger) saves no Newton-Krylov work worth having; the last two corrections,
    # whose linear tolerance the forcing terms set from the residual history, dominate the count
    far = [sol.jvp_evaluations for kind, eps, sol in rows[1:] if eps >= 0.01]
    assert min(far) >= 0.9 * zero.jvp_evaluations

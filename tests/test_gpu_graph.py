"""The Newton and EOM operator chains inside a HIP graph.

Every launch of a Jacobian product (tangent unpack, tangent atoms, JVP program, pack) and of a
sigma product (input atoms, sigma program) goes to the caller's stream with no allocation, copy
or synchronisation inside, so a caller can capture either chain in a HIP graph (here through
torch's graph capture on the library's stream) and replay it.  The test replays each captured
chain several times with eager work of the same plan between the replays (other inputs, other
precision, residuals that reload the amplitude atoms) and checks every replay bit for bit against
the eager product, and prints the time per product of eager launches against graph replay.
"""

from __future__ import annotations

import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _timed(fn, stream, reps=100):
    stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    stream.synchronize()
    return (time.perf_counter() - t0) / reps


@pytest.mark.parametrize("bits", [64, 32])
def test_jvp_chain_replays_in_a_hip_graph(torch_gpu, bits):
    from c3lindblad_amd.lowering import factorized_cc_problem
    from c3lindblad_amd.static import ComponentStaticProblem
    from tests.conftest import random_model

    print(f"case hip graph jvp fp{bits}", flush=True)
    torch = torch_gpu
    p = factorized_cc_problem(random_model(7, 1.3, 31))
    dev = ComponentStaticProblem.compile(p).device
    rng = np.random.default_rng(2)
    x = torch.from_numpy(0.05 * rng.normal(size=p.num_variables)).cuda()
    r = torch.empty_like(x)
    dev.residual_real(x, r)
    dx = torch.from_numpy(rng.normal(size=p.num_variables)).cuda()
    other = torch.from_numpy(rng.normal(size=p.num_variables)).cuda()
    out = torch.empty_like(x)
    stream = torch.cuda.Stream()
    dev.plan.set_precision("D", "jvp", bits)
    try:
        with torch.cuda.stream(stream):
            for _ in range(2):  # first launches build the program images outside the capture
                dev.jvp_real(dx, out)
            want = out.clone()
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph, stream=stream):
                dev.jvp_real(dx, out)
            for k in range(4):
                scratch = torch.empty_like(x)
                dev.jvp_real(other, scratch)  # eager work of the same plan between replays
                out.zero_()
                graph.replay()
                stream.synchronize()
                assert torch.equal(out, want), f"replay {k}"
            eager = _timed(lambda: dev.jvp_real(dx, out), stream)
            replay = _timed(graph.replay, stream)
    finally:
        dev.plan.set_precision("D", "jvp", 64)
    print(f"  JVP n=7 fp{bits}: eager {eager * 1e6:.1f} us, graph replay {replay * 1e6:.1f} us", flush=True)


def test_sigma_chain_replays_in_a_hip_graph(torch_gpu):
    from c3lindblad_amd.eom import ComponentEOMProblem, component_eom_context, eom_ccsd_problem
    from c3lindblad_amd.lowering import factorized_cc_problem
    from c3lindblad_amd.static import ComponentStaticProblem
    from tests.conftest import random_model

    print("case hip graph sigma", flush=True)
    torch = torch_gpu
    p = factorized_cc_problem(random_model(7, 1.3, 37))
    sp = ComponentStaticProblem.compile(p)
    rng = np.random.default_rng(3)
    t = p.tilde.unpack(0.05 * rng.normal(size=p.num_variables))
    ctx = component_eom_context(ComponentEOMProblem.compile(eom_ccsd_problem(p), sp, -1), t)
    dev = ctx.device
    P = ctx.problem.dimension
    v = torch.from_numpy(rng.normal(size=P) + 1j * rng.normal(size=P)).cuda()
    w = torch.from_numpy(rng.normal(size=P) + 1j * rng.normal(size=P)).cuda()
    out = torch.empty_like(v)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        for _ in range(2):
            dev.sigma(v, out)
        want = out.clone()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=stream):
            dev.sigma(v, out)
        for k in range(4):
            scratch = torch.empty_like(v)
            dev.sigma(w, scratch)
            out.zero_()
            graph.replay()
            stream.synchronize()
            assert torch.equal(out, want), f"replay {k}"
        eager = _timed(lambda: dev.sigma(v, out), stream)
        replay = _timed(graph.replay, stream)
    print(f"  sigma n=7: eager {eager * 1e6:.1f} us, graph replay {replay * 1e6:.1f} us", flush=True)

"""Device (HIP) parity against the CPU oracle.

Bar: FP64 results within 1e-12 relative (max-norm, relative to the output
scale) of the polynomial oracle for residual, JVP and sigma at small sizes and
of the NumPy plan port at larger sizes; Newton root with packed residual
<= 1e-10; G^R within 1e-8 relative of the FP64 oracle.
"""

from __future__ import annotations

import numpy as np
import pytest

from tests.conftest import random_model

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _static(n, U, seed):
    from c3lindblad_amd.lowering import factorized_cc_problem
    from c3lindblad_amd.static import ComponentStaticProblem

    p = factorized_cc_problem(random_model(n, U, seed))
    return p, ComponentStaticProblem.compile(p)


@pytest.mark.parametrize("n,U,seed", [(1, 1.3, 1), (2, 0.9, 2), (3, 1.2, 3)])
def test_residual_and_jvp_match_polynomial_oracle(torch_gpu, n, U, seed):
    from c3lindblad_amd.static import component_static_jvp, component_static_residual
    from oracle import polynomial as poly

    print(f"case residual/jvp n={n} U={U} seed={seed}", flush=True)
    p, sp = _static(n, U, seed)
    rng = np.random.default_rng(11)
    t = p.tilde.unpack(0.05 * rng.normal(size=p.num_variables))
    dt = p.tilde.unpack(rng.normal(size=p.num_variables))
    R, ref = poly.residual(p, t)
    J = poly.jvp(p, t, dt)
    got = component_static_residual(sp, t)
    assert _rel(got.retained, R) < RTOL
    assert abs(got.reference - ref) < RTOL * max(1.0, abs(ref))
    assert _rel(component_static_jvp(sp, t, dt).retained, J) < RTOL


@pytest.mark.parametrize("n", [7, 17])
def test_residual_and_jvp_match_port(torch_gpu, n):
    """n = 17: rank-4 contractions exceed one workgroup's tree, so the first contraction
    level runs as its own launch and the later levels as contraction trees."""

    from c3lindblad_amd.static import component_static_jvp, component_static_residual
    from oracle.port import PortPlan

    print(f"case residual/jvp port n={n}", flush=True)
    p, sp = _static(n, 2.0, 5)
    rng = np.random.default_rng(12)
    t = p.tilde.unpack(0.02 * rng.normal(size=p.num_variables))
    dt = p.tilde.unpack(rng.normal(size=p.num_variables))
    port = PortPlan(sp.lowered)
    port.load(0, t)
    R = port.apply(0, np.zeros(p.num_amplitudes, complex), sp.diagonal, t)
    port.load(1, dt)
    J = port.apply(2, np.zeros(p.num_amplitudes, complex), sp.diagonal, dt)
    assert _rel(component_static_residual(sp, t).retained, R) < RTOL
    assert _rel(component_static_jvp(sp, t, dt).retained, J) < RTOL


@pytest.mark.parametrize("n,charge", [(2, 1), (2, -1), (3, 1), (3, -1)])
def test_sigma_matches_polynomial_oracle(torch_gpu, n, charge):
    from c3lindblad_amd.eom import ComponentEOMProblem, component_eom_context, component_eom_sigma, eom_ccsd_problem
    from oracle import polynomial as poly

    print(f"case sigma n={n} charge={charge}", flush=True)
    p, sp = _static(n, 1.1, 20 + n)
    rng = np.random.default_rng(13)
    t = p.tilde.unpack(0.05 * rng.normal(size=p.num_variables))
    ep = ComponentEOMProblem.compile(eom_ccsd_problem(p), sp, charge)
    ctx = component_eom_context(ep, t)
    x = rng.normal(size=ep.dimension) + 1j * rng.normal(size=ep.dimension)
    want = poly.sigma(p, t, poly.coordinate_masks(ep.coordinates), x)
    assert _rel(component_eom_sigma(ctx, x).retained, want) < RTOL


def test_sigma_matches_port_n8(torch_gpu):
    from c3lindblad_amd.eom import ComponentEOMProblem, component_eom_context, component_eom_sigma, eom_ccsd_problem
    from oracle.port import PortPlan

    print("case sigma port n=8", flush=True)
    p, sp = _static(8, 3.0, 9)
    rng = np.random.default_rng(14)
    t = p.tilde.unpack(0.02 * rng.normal(size=p.num_variables))
    ep = ComponentEOMProblem.compile(eom_ccsd_problem(p), sp, 1)
    ctx = component_eom_context(ep, t)
    x = rng.normal(size=ep.dimension) + 1j * rng.normal(size=ep.dimension)
    port = PortPlan(ep.lowered)
    port.load(0, t)
    port.load(2, x)
    want = port.apply(3, np.zeros(ep.dimension, complex), ep.diagonal, x)
    assert _rel(component_eom_sigma(ctx, x).retained, want) < RTOL


def test_tilde_round_trip_and_krylov_kernels(torch_gpu):
    import torch

    from c3lindblad_amd.device import DeviceTilde, Reducer

    print("case tilde/krylov kernels", flush=True)
    p, _ = _static(4, 1.0, 4)
    tl = DeviceTilde(p.tilde)
    rng = np.random.default_rng(5)
    x = rng.normal(size=p.num_variables)
    xd = torch.from_numpy(x).cuda()
    td = torch.empty(p.num_amplitudes, dtype=torch.complex128, device="cuda")
    tl.unpack(xd, td)
    assert np.max(np.abs(td.cpu().numpy() - p.tilde.unpack(x))) < 1e-15
    back = torch.empty_like(xd)
    tl.pack(td, back)
    assert np.max(np.abs(back.cpu().numpy() - x)) < 1e-14
    # complex multi-dot / multi-axpy over a ragged dimension (not a chunk multiple)
    dim, cols = 20011, 37
    V = rng.normal(size=(cols, dim)) + 1j * rng.normal(size=(cols, dim))
    w = rng.normal(size=dim) + 1j * rng.normal(size=dim)
    Vd, wd = torch.from_numpy(V).cuda(), torch.from_numpy(w).cuda()
    red = Reducer(dim, cols, complex_=True)
    out = torch.empty(cols, dtype=torch.complex128, device="cuda")
    red.dot(Vd, cols, wd, out)
    assert _rel(out.cpu().numpy(), V.conj() @ w) < 1e-13
    red.dot(Vd, cols, wd, out, conjugate=False)
    assert _rel(out.cpu().numpy(), V @ w) < 1e-13
    alpha = rng.normal(size=cols) + 1j * rng.normal(size=cols)
    red.axpy(Vd, cols, torch.from_numpy(alpha).cuda(), wd, -0.5)
    assert _rel(wd.cpu().numpy(), w - 0.5 * alpha @ V) < 1e-13
    rr = Reducer(dim, cols, complex_=False)
    Vr, wr = torch.from_numpy(V.real.copy()).cuda(), torch.from_numpy(w.real.copy()).cuda()
    outr = torch.empty(cols, dtype=torch.float64, device="cuda")
    rr.dot(Vr, cols, wr, outr)
    assert _rel(outr.cpu().numpy(), V.real @ w.real) < 1e-13


def test_shifted_triangular_green_kernel(torch_gpu):
    import scipy.linalg

    from c3lindblad_amd.response import shifted_green

    print("case shifted triangular", flush=True)
    rng = np.random.default_rng(6)
    m = 97
    H = np.triu(rng.normal(size=(m, m)) + 1j * rng.normal(size=(m, m)), -1)
    T, Q = scipy.linalg.schur(H, output="complex")
    r = rng.normal(size=m) + 1j * rng.normal(size=m)
    w = rng.normal(size=m) + 1j * rng.normal(size=m)
    q = rng.normal(size=m) + 1j * rng.normal(size=m)
    z = np.linspace(-3, 3, 301) + 0.5j
    g, last = shifted_green(T, r, w, q, z)
    for k in (0, 150, 300):
        y = np.linalg.solve(z[k] * np.eye(m) - T, r)
        assert abs(g[k] - w @ y) < 1e-11 * max(1, abs(w @ y))
        assert abs(last[k] - q @ y) < 1e-11 * max(1, abs(q @ y))


def test_newton_on_device_matches_cpu_oracle_root(torch_gpu):
    from c3lindblad_amd.configs import readme_model
    from c3lindblad_amd.lowering import factorized_cc_problem
    from c3lindblad_amd.solver import solve_component_ccsd
    from c3lindblad_amd.static import ComponentStaticProblem
    from oracle import polynomial as poly
    from oracle import solvers

    print("case newton readme model", flush=True)
    p = factorized_cc_problem(readme_model())
    sp = ComponentStaticProblem.compile(p)
    sol = solve_component_ccsd(sp)
    assert sol.converged and sol.packed_residual_norm <= 1e-10
    masks = poly.coordinate_masks(p.coordinates)
    _, t_cpu, packed, _, _ = solvers.solve_ccsd(
        p.tilde, p.rapidity_sums(), lambda t: poly.residual(p, t, masks)[0], lambda t: (lambda d: poly.jvp(p, t, d, masks))
    )
    assert packed <= 1e-10
    # both roots satisfy the same equations; they agree to the conditioning of the Jacobian
    assert _rel(sol.amplitudes, t_cpu) < 1e-8
    R, _ = poly.residual(p, sol.amplitudes, masks)
    assert np.linalg.norm(p.tilde.pack(R, check=False)) <= 1e-10


def test_readme_green_function_matches_dense_oracle(torch_gpu):
    from c3lindblad_amd.configs import readme_grid, readme_model
    from c3lindblad_amd.eom import (
        ComponentEOMProblem,
        ComponentSigmaOperator,
        component_eom_context,
        component_eom_sources,
        eom_ccsd_problem,
    )
    from c3lindblad_amd.lowering import factorized_cc_problem
    from c3lindblad_amd.response import component_retarded_green
    from c3lindblad_amd.solver import solve_component_ccsd
    from c3lindblad_amd.static import ComponentStaticProblem
    from oracle import polynomial as poly
    from oracle import solvers

    print("case readme G^R", flush=True)
    p = factorized_cc_problem(readme_model())
    sp = ComponentStaticProblem.compile(p)
    sol = solve_component_ccsd(sp)
    omega = readme_grid()
    ep = ComponentEOMProblem.compile(eom_ccsd_problem(p), sp, 1)
    ctx = component_eom_context(ep, sol.amplitudes)
    sources = component_eom_sources(ctx)
    result = component_retarded_green(ComponentSigmaOperator(ctx), sources, omega, 0.0)
    masks = poly.coordinate_masks(ep.coordinates)
    smasks = poly.coordinate_masks(p.coordinates)
    A = solvers.dense_matrix(lambda v: poly.sigma(p, sol.amplitudes, masks, v, smasks), ep.dimension)
    add, rem = solvers.retarded_green_dense(A, sources.addition, sources.removal, sources.sink, omega, 0.0)
    assert _rel(result.retarded, add + rem) < 1e-8
    assert np.max(result.addition_relative_residuals) <= 1e-9


def test_readme_solve_matches_the_reference_golden_vectors(torch_gpu):
    """The device path (Newton-Krylov CCSD and the shared-basis Arnoldi resolvent) on the README
    quick-start model against the reference's own NumPy results stored in
    ``tests/golden/config1_reference_vectors.npz``: amplitudes in the reference's gauge and order
    within 1e-10, charge +1 and -1 G^R on the 81 frequencies and the occupations within 1e-8
    relative."""

    from pathlib import Path

    from c3lindblad_amd.configs import readme_model
    from c3lindblad_amd.eom import (
        ComponentEOMProblem,
        ComponentSigmaOperator,
        component_eom_context,
        component_eom_sources,
        eom_ccsd_problem,
    )
    from c3lindblad_amd.lowering import factorized_cc_problem
    from c3lindblad_amd.response import component_retarded_green
    from c3lindblad_amd.solver import solve_component_ccsd
    from c3lindblad_amd.static import ComponentStaticProblem

    print("case readme vs reference golden vectors", flush=True)
    g = np.load(Path(__file__).parent / "golden" / "config1_reference_vectors.npz")
    p = factorized_cc_problem(readme_model())
    sp = ComponentStaticProblem.compile(p)
    sol = solve_component_ccsd(sp)
    assert sol.converged and sol.packed_residual_norm <= 1e-10
    assert np.max(np.abs(sol.amplitudes - g["amplitudes"])) < 1e-10 * np.max(np.abs(g["amplitudes"]))
    for charge, key, tag in ((1, "retarded", "plus"), (-1, "retarded_minus", "minus")):
        ep = ComponentEOMProblem.compile(eom_ccsd_problem(p), sp, charge)
        ctx = component_eom_context(ep, sol.amplitudes)
        result = component_retarded_green(ComponentSigmaOperator(ctx), component_eom_sources(ctx), g["omega"], 0.0)
        assert _rel(result.retarded, g[key]) < 1e-8, charge
        assert _rel(result.addition, g[f"addition_{tag}"]) < 1e-8, charge  # each channel, not only the sum
        assert _rel(result.removal, g[f"removal_{tag}"]) < 1e-8, charge
    from c3lindblad_amd.observables import factorized_ccsd_covariance

    assert _rel(factorized_ccsd_covariance(sp, sol.amplitudes), g["covariance"]) < 1e-8


@pytest.mark.parametrize("name", ["siam6", "multiorbital6"])
def test_family_solves_match_the_reference_golden_vectors(torch_gpu, name):
    """The default mixed-precision device solve (FP32 JVPs inside GMRES) of two n = 6 models of the
    benchmark families against the reference's NumPy results (``tests/golden/*_reference_vectors.npz``):
    amplitudes within 1e-9 relative, charge +1 G^R on 61 frequencies within 1e-8 relative."""

    from pathlib import Path

    from c3lindblad_amd.configs import multiorbital_model, siam_model
    from c3lindblad_amd.eom import (
        ComponentEOMProblem,
        ComponentSigmaOperator,
        component_eom_context,
        component_eom_sources,
        eom_ccsd_problem,
    )
    from c3lindblad_amd.lowering import factorized_cc_problem
    from c3lindblad_amd.response import component_retarded_green
    from c3lindblad_amd.solver import NewtonKrylovPolicy, solve_component_ccsd
    from c3lindblad_amd.static import ComponentStaticProblem

    print(f"case {name} vs reference golden vectors", flush=True)
    model = siam_model(5, 4.0, seed=3) if name == "siam6" else multiorbital_model(2, 2, seed=3)
    g = np.load(Path(__file__).parent / "golden" / f"{name}_reference_vectors.npz")
    p = factorized_cc_problem(model)
    sp = ComponentStaticProblem.compile(p)
    policy = NewtonKrylovPolicy.mixed_precision(forcing_strategy="eisenstat_walker", gmres_restart=150)
    sol = solve_component_ccsd(sp, policy=policy)
    assert sol.converged and sol.packed_residual_norm <= 1e-10
    assert np.max(np.abs(sol.amplitudes - g["amplitudes"])) < 1e-9 * np.max(np.abs(g["amplitudes"]))
    ep = ComponentEOMProblem.compile(eom_ccsd_problem(p), sp, 1)
    ctx = component_eom_context(ep, sol.amplitudes)
    result = component_retarded_green(ComponentSigmaOperator(ctx), component_eom_sources(ctx), g["omega"], 0.0)
    assert _rel(result.retarded, g["retarded"]) < 1e-8


def test_single_site_green_function_matches_lindblad_ed(torch_gpu):
    from c3lindblad_amd import (
        ComponentEOMProblem,
        ComponentSigmaOperator,
        ComponentStaticProblem,
        FermionicLindbladModel,
        component_eom_context,
        component_eom_sources,
        component_retarded_green,
        eom_ccsd_problem,
        factorized_cc_problem,
        solve_component_ccsd,
    )
    from oracle import fock

    print("case single site vs ED", flush=True)
    model = FermionicLindbladModel(E=[[-0.45]], gamma_loss=[[0.35]], gamma_gain=[[0.15]], U=1.3)
    p = factorized_cc_problem(model)
    sp = ComponentStaticProblem.compile(p)
    sol = solve_component_ccsd(sp)
    assert sol.converged
    ep = ComponentEOMProblem.compile(eom_ccsd_problem(p), sp, 1)
    ctx = component_eom_context(ep, sol.amplitudes)
    omega = np.linspace(-3.0, 3.0, 31)
    result = component_retarded_green(ComponentSigmaOperator(ctx), component_eom_sources(ctx), omega, 0.0)
    want = fock.exact_retarded_green(model, omega, 0.0)
    assert _rel(result.retarded, want) < 1e-8


@pytest.mark.parametrize("restart,cycles,rtol,atol", [(40, 12, 1e-6, 0.0), (5, 8, 1e-9, 0.0), (6, 10, 1e-3, 1e-7)])
def test_native_gmres_matches_scipy(torch_gpu, restart, cycles, rtol, atol):
    """c3l_gmres against scipy.sparse.linalg.gmres on the oracle Jacobian: same inner
    iteration count, same convergence flag, same step to rounding."""

    from scipy.sparse.linalg import LinearOperator, gmres

    from c3lindblad_amd.solver import _DeviceNewton
    from oracle import polynomial as poly
    from oracle import solvers

    print(f"case gmres restart={restart} cycles={cycles} rtol={rtol} atol={atol}", flush=True)
    p, sp = _static(3, 1.2, 3)
    torch = torch_gpu
    rng = np.random.default_rng(7)
    D = p.tilde.real_size
    x = 0.05 * rng.normal(size=D)
    state = _DeviceNewton(sp, restart)
    xd = torch.from_numpy(x).cuda()
    norm = state.residual(xd, state.r)
    b = (-state.r).contiguous()
    step = torch.empty_like(b)
    info, its, true = state.gmres(b, step, rtol, atol, restart, cycles)
    w = state.w.cpu().numpy()

    masks = poly.coordinate_masks(p.coordinates)
    t = p.tilde.unpack(x)
    r = p.tilde.pack(poly.residual(p, t, masks)[0], check=False)
    assert _rel(-b.cpu().numpy(), r) < RTOL
    diag = solvers.preconditioner_diagonal(p.rapidity_sums())
    A = LinearOperator((D, D), matvec=lambda v: p.tilde.pack(poly.jvp(p, t, p.tilde.unpack(v), masks), check=False),
                       dtype=np.float64)
    M = LinearOperator((D, D), matvec=lambda v: p.tilde.pack(p.tilde.unpack(v) / diag, check=False), dtype=np.float64)
    seen = []
    want, want_info = gmres(A, -r, M=M, rtol=rtol, atol=atol, restart=restart, maxiter=cycles,
                            callback=lambda v: seen.append(v), callback_type="pr_norm")
    assert its == len(seen)
    assert (info == 0) == (want_info == 0)
    step_err = _rel(step.cpu().numpy(), want)
    assert step_err < 1e-7
    # w is the true linear residual b - J step (a cancellation: compare on the scale of b)
    w_err = float(np.max(np.abs(w - (-r - A.matvec(step.cpu().numpy())))) / np.max(np.abs(r)))
    assert w_err < 1e-12
    assert abs(true - np.linalg.norm(w) / norm) <= 1e-12
    assert (info == 0) == (true * norm <= max(rtol * norm, atol))


def test_native_gmres_with_fp32_basis_meets_the_tolerance(torch_gpu):
    """FP32-stored Krylov basis (FP64 dots, coefficients and updates): the true residual
    meets rtol 1e-6 and the step agrees with SciPy's to the tolerance's scale."""

    from scipy.sparse.linalg import LinearOperator, gmres

    from c3lindblad_amd.solver import _DeviceNewton
    from oracle import polynomial as poly
    from oracle import solvers

    print("case gmres fp32 basis", flush=True)
    p, sp = _static(3, 1.2, 3)
    torch = torch_gpu
    rng = np.random.default_rng(7)
    D = p.tilde.real_size
    x = 0.05 * rng.normal(size=D)
    state = _DeviceNewton(sp, 40, basis_bits=32)
    assert state.V.dtype == torch.float32
    norm = state.residual(torch.from_numpy(x).cuda(), state.r)
    b = (-state.r).contiguous()
    step = torch.empty_like(b)
    info, its, true = state.gmres(b, step, 1e-6, 0.0, 40, 12)
    assert info == 0 and true <= 1e-6
    masks = poly.coordinate_masks(p.coordinates)
    t = p.tilde.unpack(x)
    r = p.tilde.pack(poly.residual(p, t, masks)[0], check=False)
    diag = solvers.preconditioner_diagonal(p.rapidity_sums())
    A = LinearOperator((D, D), matvec=lambda v: p.tilde.pack(poly.jvp(p, t, p.tilde.unpack(v), masks), check=False),
                       dtype=np.float64)
    M = LinearOperator((D, D), matvec=lambda v: p.tilde.pack(p.tilde.unpack(v) / diag, check=False), dtype=np.float64)
    want, want_info = gmres(A, -r, M=M, rtol=1e-10, atol=0.0, restart=200, maxiter=5)
    assert want_info == 0
    assert _rel(step.cpu().numpy(), want) < 1e-4
    # the returned residual is the exact FP64 one of the returned step
    w = state.w.cpu().numpy()
    assert float(np.max(np.abs(w - (-r - A.matvec(step.cpu().numpy())))) / np.max(np.abs(r))) < 1e-12


def test_eisenstat_walker_newton_reaches_legacy_root(torch_gpu):
    from c3lindblad_amd.solver import NewtonKrylovPolicy, solve_component_ccsd

    print("case newton eisenstat-walker", flush=True)
    p, sp = _static(3, 1.2, 3)
    legacy = solve_component_ccsd(sp)
    ew = solve_component_ccsd(sp, policy=NewtonKrylovPolicy(forcing_strategy="eisenstat_walker"))
    assert legacy.converged and ew.converged
    assert ew.packed_residual_norm <= 1e-10
    assert all(step.damping > 0.0 for step in ew.history)
    assert _rel(ew.amplitudes, legacy.amplitudes) < 1e-8


@pytest.mark.parametrize("n,U,seed", [(2, 0.9, 2), (3, 1.2, 3)])
def test_device_observables_match_oracle(torch_gpu, n, U, seed):
    from c3lindblad_amd.observables import factorized_ccsd_covariance, two_body_expectations
    from oracle import polynomial as poly

    print(f"case observables n={n}", flush=True)
    p, sp = _static(n, U, seed)
    rng = np.random.default_rng(seed + 30)
    t = p.tilde.unpack(0.1 * rng.normal(size=p.num_variables))
    got = factorized_ccsd_covariance(sp, t)
    assert _rel(got, poly.covariance(p, t)) < RTOL
    N = p.model.n_spin_orbitals
    quartets = rng.integers(0, N, size=(8, 4))
    got4 = two_body_expectations(sp, t, quartets)
    from c3lindblad_amd.lowering import physical_mode_operator

    ref = p.reference
    op = {(o, d): physical_mode_operator(ref, *ref.model.orbital_site_spin(o), "c_dagger" if d else "c").components
          for o in range(N) for d in (0, 1)}
    prefix = poly.low_rank_prefix(poly.coordinate_masks(p.coordinates), t)
    want4 = np.array([poly.expectation(prefix, (op[a, 1], op[b, 0], op[c, 1], op[d, 0])) for a, b, c, d in quartets])
    assert np.max(np.abs(got4 - want4)) < RTOL * max(1.0, np.max(np.abs(want4)))


def test_single_site_device_occupations_equal_lindblad_ed(torch_gpu):
    from c3lindblad_amd import (
        ComponentStaticProblem,
        FermionicLindbladModel,
        factorized_cc_problem,
        factorized_ccsd_covariance,
        solve_component_ccsd,
    )
    from oracle import fock

    print("case single site occupations vs ED", flush=True)
    model = FermionicLindbladModel(E=[[-0.45]], gamma_loss=[[0.35]], gamma_gain=[[0.15]], U=1.3)
    sp = ComponentStaticProblem.compile(factorized_cc_problem(model))
    sol = solve_component_ccsd(sp)
    got = factorized_ccsd_covariance(sp, sol.amplitudes)
    assert np.max(np.abs(got - fock.exact_covariance(model))) < 1e-10


def test_fp32_jvp_matches_fp64_jvp(torch_gpu):
    """Mixed-precision JVP: complex64 atoms with FP32 accumulation against the FP64
    JVP of the same plan, relative (max-norm) error <= 2e-6 at n = 7."""

    print("case fp32 jvp n=7", flush=True)
    p, sp = _static(7, 1.3, 7)
    dev = sp.device
    torch = torch_gpu
    rng = np.random.default_rng(3)
    x = torch.from_numpy(0.05 * rng.normal(size=p.num_variables)).cuda()
    r = torch.empty_like(x)
    dev.residual_real(x, r)
    dx = torch.from_numpy(rng.normal(size=p.num_variables)).cuda()
    want = torch.empty_like(x)
    dev.jvp_real(dx, want)
    got = torch.empty_like(x)
    dev.plan.set_precision("D", "jvp", 32)
    try:
        dev.jvp_real(dx, got)
        dev.jvp_real(dx, got)  # second call: narrowed atoms are reused
    finally:
        dev.plan.set_precision("D", "jvp", 64)
    g, w = got.cpu().numpy(), want.cpu().numpy()
    err = _rel(g, w)
    assert 1e-12 < err < 2e-6  # FP32 rounding is visible, and bounded
    again = torch.empty_like(x)
    dev.jvp_real(dx, again)
    assert _rel(again.cpu().numpy(), w) < RTOL  # back to FP64


def _matrix_core_case(name):
    from c3lindblad_amd.configs import multiorbital_model
    from c3lindblad_amd.lowering import factorized_cc_problem
    from c3lindblad_amd.static import ComponentStaticProblem

    if name.startswith("random"):
        return _static(int(name.split("_")[1]), 1.7, 9)
    model = multiorbital_model(3, 2, seed=2)  # three Kanamori orbitals with V_pqrs scatter
    p = factorized_cc_problem(model)
    return p, ComponentStaticProblem.compile(p)


@pytest.mark.parametrize("case", ["random_3", "random_7", "random_17", "kanamori_3x2"])
def test_fp32_matrix_core_jvp(torch_gpu, case):
    """FP32 JVP with the rank-4 two-atom blocks on the matrix cores (16x4x4x4 output tiles; n = 3,
    7 and 17 leave partial tiles; at n = 17 the (2,2) blocks with antisymmetric coordinates run on
    tiles too) against the FP32 gather kernels and the FP64 JVP: both FP32 variants within 2e-6
    (max-norm relative) of FP64 and within 2e-6 of each other, and the matrix-core variant must
    have run (tiles > 0)."""

    print(f"case fp32 matrix-core jvp {case}", flush=True)
    p, sp = _matrix_core_case(case)
    dev = sp.device
    torch = torch_gpu
    rng = np.random.default_rng(4)
    x = torch.from_numpy(0.05 * rng.normal(size=p.num_variables)).cuda()
    r = torch.empty_like(x)
    dev.residual_real(x, r)
    dx = torch.from_numpy(rng.normal(size=p.num_variables)).cuda()
    want = torch.empty_like(x)
    dev.jvp_real(dx, want)
    got = {}
    dev.plan.set_precision("D", "jvp", 32)
    try:
        for cores in (0, 2, 2):  # the last call reuses the built tiles
            dev.plan.set_matrix_cores(cores)
            out = torch.empty_like(x)
            dev.jvp_real(dx, out)
            got[cores] = out.cpu().numpy()
        tiles = dev.plan.matrix_core_tiles("jvp")
    finally:
        dev.plan.set_precision("D", "jvp", 64)
        dev.plan.set_matrix_cores(1)
    w = want.cpu().numpy()
    e_gather, e_cores, e_pair = _rel(got[0], w), _rel(got[2], w), _rel(got[2], got[0])
    print(f"  tiles={tiles} gather={e_gather:.2e} cores={e_cores:.2e} cores-vs-gather={e_pair:.2e}", flush=True)
    assert tiles > 0
    assert e_gather < 2e-6 and e_cores < 2e-6 and e_pair < 2e-6


@pytest.mark.parametrize("n,seed", [(7, 23), (17, 29)])
def test_autotuned_program_shapes_keep_the_results(torch_gpu, n, seed):
    """Launch-shape autotuner: residual, FP64 and FP32 JVP and sigma after tuning against the
    same programs on the default shapes (FP64 within 1e-13 relative: only the summation order of
    a coordinate's entries may change, and bit for bit equal when the tuner kept the default lane
    split; FP32 within 2e-6 of the FP64 JVP), every tuned time at most the default one, and every
    chosen shape one of the candidates."""

    from c3lindblad_amd.eom import ComponentEOMProblem, component_eom_context, component_eom_sigma, eom_ccsd_problem
    from c3lindblad_amd.tuning import autotune_sigma, autotune_static

    print(f"case autotune n={n}", flush=True)
    p, sp = _static(n, 1.2, seed)  # a fresh digest: nothing tuned for it yet
    dev = sp.device
    torch = torch_gpu
    rng = np.random.default_rng(seed)
    x = torch.from_numpy(0.05 * rng.normal(size=p.num_variables)).cuda()
    dx = torch.from_numpy(rng.normal(size=p.num_variables)).cuda()
    ep = ComponentEOMProblem.compile(eom_ccsd_problem(p), sp, 1)
    t = p.tilde.unpack(0.05 * rng.normal(size=p.num_variables))
    v = rng.normal(size=ep.dimension) + 1j * rng.normal(size=ep.dimension)

    def evaluate():
        res, j64, j32 = (torch.empty_like(x) for _ in range(3))
        dev.residual_real(x, res)
        dev.jvp_real(dx, j64)
        dev.plan.set_precision("D", "jvp", 32)
        try:
            dev.jvp_real(dx, j32)
        finally:
            dev.plan.set_precision("D", "jvp", 64)
        sigma = component_eom_sigma(component_eom_context(ep, t), v).retained
        return [a.cpu().numpy() for a in (res, j64, j32)] + [sigma]

    before = evaluate()
    records = autotune_static(sp, 64) + autotune_static(sp, 32) + [autotune_sigma(ep)]
    after = evaluate()
    paths = [r for r in records if r["program"] == "residual_path"]
    records = [r for r in records if r["program"] != "residual_path"]
    assert len(paths) == 2 and all(r["choice"] in ("program", "quadrature", "simpson") for r in paths)
    print(f"  residual path: {paths[-1]}", flush=True)
    for r in records:
        print(f"  {r['program']}: {r['ms_default'] * 1e3:.1f} -> {r['ms_tuned'] * 1e3:.1f} us {r['shape']}", flush=True)
        assert r["ms_tuned"] <= r["ms_default"]
        sh = r["shape"]
        assert sh["lanes_max_log2"] in (0, 1, 2) and sh["coordinates_per_thread"] in (1, 2)
        assert sh["lane_split_workgroups"] in (256, 512, 1024, 2048) and sh["xcd_slicing"] in (0, 1)
        assert sh["coordinate_workgroups"] in (256, 512, 1024, 2048)
        assert sh["matrix_core_min_entries"] in (1, 32, 128, 2**31 - 1)
    res0, j640, j320, s0 = before
    res1, j641, j321, s1 = after
    lanes = records[0]["shape"]  # residual, FP64
    if lanes["lanes_max_log2"] == 2 and lanes["lane_split_workgroups"] == 512 and sp.device.residual_path == "program":
        assert np.array_equal(res1, res0)  # order-preserving knobs only
    assert _rel(res1, res0) < 1e-13
    assert _rel(j641, j640) < 1e-13
    assert _rel(s1, s0) < 1e-13
    assert _rel(j321, j640) < 2e-6 and _rel(j320, j640) < 2e-6


@pytest.mark.parametrize("case,path", [("random_3", "quadrature"), ("random_7", "quadrature"),
                                       ("kanamori_3x2", "quadrature"), ("random_3", "simpson"),
                                       ("random_7", "simpson"), ("kanamori_3x2", "simpson")])
def test_residual_by_quadrature_matches_the_residual_program(torch_gpu, case, path):
    """Residual as R(0) + sum_i w_i J(s_i t) t at the Gauss-Legendre nodes of the residual's
    degree (two nodes for degree 4), or by Simpson's rule (JVPs at s = 1/2 and 1 plus the
    residual_low program), against the residual program (1e-13 relative), against the
    polynomial oracle at n = 3 (1e-12), entered with the JVP in FP32 (the FP32 state must be back
    afterwards), and a mixed-precision Newton solve on that path reaching the root of
    the program path (packed residual <= 1e-10, amplitudes within 1e-8)."""

    from c3lindblad_amd.solver import NewtonKrylovPolicy, solve_component_ccsd
    from oracle import polynomial as poly

    print(f"case {path} residual {case}", flush=True)
    p, sp = _matrix_core_case(case)
    dev = sp.device
    torch = torch_gpu
    assert sp.residual_degree == 4 and len(dev.quadrature_nodes()[0]) == 2
    rng = np.random.default_rng(6)
    x = torch.from_numpy(0.05 * rng.normal(size=p.num_variables)).cuda()
    got = {}
    for which in ("program", path, path):  # the second call reuses the cached constant parts
        dev.residual_path = which
        r = torch.empty_like(x)
        dev.residual_real(x, r)
        got[which] = (r.cpu().numpy(), dev.out.cpu().numpy(), complex(dev.ref.item()))
    assert _rel(got[path][0], got["program"][0]) < 1e-13
    assert abs(got[path][2] - got["program"][2]) <= 1e-13 * max(1.0, abs(got["program"][2]))
    if case == "random_3":
        want, _ = poly.residual(p, p.tilde.unpack(x.cpu().numpy()))
        assert _rel(got[path][1], want) < 1e-12
    # entered while the JVP program is FP32: quadrature JVPs run in FP64, FP32 comes back
    dx = torch.from_numpy(rng.normal(size=p.num_variables)).cuda()
    j64 = torch.empty_like(x)
    dev.residual_path = "program"
    dev.residual_real(x, torch.empty_like(x))
    dev.jvp_real(dx, j64)
    dev.plan.set_precision("D", "jvp", 32)
    try:
        dev.residual_path = path
        r = torch.empty_like(x)
        dev.residual_real(x, r)
        assert dev.plan.precision("jvp") == 32
        j32 = torch.empty_like(x)
        dev.jvp_real(dx, j32)
    finally:
        dev.plan.set_precision("D", "jvp", 64)
    assert _rel(r.cpu().numpy(), got["program"][0]) < 1e-13
    e = _rel(j32.cpu().numpy(), j64.cpu().numpy())
    assert 1e-12 < e < 2e-6  # still FP32 after the residual, at the same amplitudes
    policy = NewtonKrylovPolicy.mixed_precision(forcing_strategy="eisenstat_walker", gmres_restart=150)
    dev.residual_path = "program"
    ref = solve_component_ccsd(sp, policy=policy)
    dev.residual_path = path
    try:
        sol = solve_component_ccsd(sp, policy=policy)
    finally:
        dev.residual_path = "program"
    print(f"  program path: converged={ref.converged} |pack R|={ref.packed_residual_norm:.2e} jvps={ref.jvp_evaluations}; "
          f"{path} path: converged={sol.converged} |pack R|={sol.packed_residual_norm:.2e} "
          f"jvps={sol.jvp_evaluations}", flush=True)
    assert ref.converged and sol.converged and sol.packed_residual_norm <= 1e-10
    scale = max(1.0, np.max(np.abs(ref.amplitudes)))
    assert np.max(np.abs(sol.amplitudes - ref.amplitudes)) < 1e-8 * scale


def test_mixed_precision_newton_reaches_fp64_root(torch_gpu):
    from c3lindblad_amd.solver import NewtonKrylovPolicy, solve_component_ccsd

    print("case mixed precision newton", flush=True)
    p, sp = _static(4, 1.6, 11)
    ref = solve_component_ccsd(sp)
    for policy in (NewtonKrylovPolicy.mixed_precision(),
                   NewtonKrylovPolicy.mixed_precision(forcing_strategy="eisenstat_walker")):
        sol = solve_component_ccsd(sp, policy=policy)
        assert sol.converged and sol.packed_residual_norm <= 1e-10
        assert _rel(sol.amplitudes, ref.amplitudes) < 1e-8


def test_solve_component_shifted_matches_dense_solve(torch_gpu):
    from c3lindblad_amd import (
        ComponentEOMProblem,
        ComponentSigmaOperator,
        ComponentStaticProblem,
        component_eom_context,
        component_eom_sources,
        eom_ccsd_problem,
        factorized_cc_problem,
        solve_component_ccsd,
        solve_component_shifted,
    )
    from c3lindblad_amd.configs import readme_model
    from oracle import polynomial as poly
    from oracle import solvers

    print("case shifted solve readme", flush=True)
    p = factorized_cc_problem(readme_model())
    sp = ComponentStaticProblem.compile(p)
    sol = solve_component_ccsd(sp)
    ep = ComponentEOMProblem.compile(eom_ccsd_problem(p), sp, 1)
    ctx = component_eom_context(ep, sol.amplitudes)
    op = ComponentSigmaOperator(ctx)
    b = component_eom_sources(ctx).addition
    masks = poly.coordinate_masks(ep.coordinates)
    A = solvers.dense_matrix(lambda v: poly.sigma(p, sol.amplitudes, masks, v, poly.coordinate_masks(p.coordinates)),
                             ep.dimension)
    for z in (0.3 + 0.0j, -1.7 + 0.05j):
        out = solve_component_shifted(op, z, b)
        want = np.linalg.solve(z * np.eye(ep.dimension) - A, b)
        assert out.converged and out.true_relative_residual <= 1e-9
        assert _rel(out.solution, want) < 1e-8
        warm = solve_component_shifted(op, z, b, initial_solution=want)
        assert warm.converged and _rel(warm.solution, want) < 1e-9


def test_arnoldi_basis_is_orthonormal_and_satisfies_the_arnoldi_relation(torch_gpu):
    """Size-independent certificate of the delayed-reorthogonalisation Arnoldi driver:
    V^H V = I and sigma V_m = V_{m+1} H_m on a 300-column basis over a ragged
    dimension (P = 744, not a multiple of the reduction chunk), including the
    last column, which a step without operator completes."""

    import torch

    from c3lindblad_amd.eom import ComponentEOMProblem, component_eom_context, eom_ccsd_problem
    from c3lindblad_amd.response import ComponentSigmaOperator, arnoldi_resolvent

    print("case arnoldi relation n=8 m=300", flush=True)
    p, sp = _static(8, 3.0, 9)
    rng = np.random.default_rng(21)
    t = p.tilde.unpack(0.02 * rng.normal(size=p.num_variables))
    ep = ComponentEOMProblem.compile(eom_ccsd_problem(p), sp, 1)
    op = ComponentSigmaOperator(component_eom_context(ep, t))
    b = rng.normal(size=ep.dimension) + 1j * rng.normal(size=ep.dimension)
    sink = rng.normal(size=ep.dimension) + 1j * rng.normal(size=ep.dimension)
    proj = arnoldi_resolvent(op, b, np.array([0.5 + 0.1j]), maximum_dimension=300, residual_tolerance=1e-300,
                             minimum_dimension=300, sink=sink)
    m = proj.dimension
    assert m == 300 and not proj.breakdown
    V = proj.basis[: m + 1]
    gram = (V.conj() @ V.T).cpu().numpy()
    assert np.max(np.abs(gram - np.eye(m + 1))) < 1e-12
    AV = torch.empty(m, ep.dimension, dtype=torch.complex128, device="cuda")
    for k in range(m):
        op.matvec_device(V[k].contiguous(), AV[k])
    H = torch.from_numpy(np.ascontiguousarray(proj.hessenberg[: m + 1, :m])).cuda()
    defect = (AV.T - V.T @ H).cpu().numpy()
    assert np.max(np.abs(defect)) / np.max(np.abs(proj.hessenberg)) < 1e-12
    # first basis vector is the normalised source, untouched by the driver
    assert np.max(np.abs(V[0].cpu().numpy() - b / np.linalg.norm(b))) < 1e-15


def test_batched_arnoldi_equals_separate_bases(torch_gpu):
    """Lockstep bases on twin plans (one launch per kernel for both) reproduce the
    separate single-basis runs bit for bit, also when one basis stops long before the
    other: its source spans six eigenvectors of sigma, so its shifts converge (or the
    basis breaks down) within a few steps and it is dropped from later launches."""

    import torch

    from c3lindblad_amd.eom import ComponentEOMProblem, component_eom_context, eom_ccsd_problem
    from c3lindblad_amd.response import ComponentSigmaOperator, arnoldi_resolvent, arnoldi_resolvent_batch

    print("case batched arnoldi n=8", flush=True)
    p, sp = _static(8, 3.0, 9)
    rng = np.random.default_rng(23)
    t = p.tilde.unpack(0.02 * rng.normal(size=p.num_variables))
    ep = ComponentEOMProblem.compile(eom_ccsd_problem(p), sp, 1)
    op = ComponentSigmaOperator(component_eom_context(ep, t))
    P = ep.dimension
    eye = torch.eye(P, dtype=torch.complex128, device="cuda")
    A = torch.empty(P, P, dtype=torch.complex128, device="cuda")
    for k in range(P):
        op.matvec_device(eye[k].contiguous(), A[k])
    A = A.cpu().numpy().T
    _, vecs = np.linalg.eig(A)
    b1 = rng.normal(size=P) + 1j * rng.normal(size=P)
    b2 = vecs[:, :6] @ (rng.normal(size=6) + 1j * rng.normal(size=6))
    sink = rng.normal(size=P) + 1j * rng.normal(size=P)
    z = np.linspace(-2.0, 2.0, 33) + 0.3j
    kw = dict(maximum_dimension=400, residual_tolerance=1e-10, minimum_dimension=8)
    pair = arnoldi_resolvent_batch([op.on_slot(0), op.on_slot(1)], [b1, b2], z, sinks=[sink, sink], **kw)
    alone = [arnoldi_resolvent(op.on_slot(s), b, z, sink=sink, criterion="gmres", **kw) for s, b in ((0, b1), (1, b2))]
    print("dimensions", [q.dimension for q in pair], [q.dimension for q in alone], flush=True)
    assert pair[1].dimension < pair[0].dimension
    # the multi-shift Green values of the long basis against dense resolvents: at this Krylov
    # dimension the shift state runs through many 32-step rotation tiles and eight-column batches
    assert pair[0].dimension > 96
    want_g = np.array([sink @ np.linalg.solve(zz * np.eye(P) - A, b1) for zz in z])
    assert _rel(pair[0].green, want_g) < 1e-8
    for got, want in zip(pair, alone):
        assert got.dimension == want.dimension and got.breakdown == want.breakdown
        assert np.array_equal(got.hessenberg, want.hessenberg)
        assert np.array_equal(got.green, want.green)
        assert np.array_equal(got.shift_relative_residuals, want.shift_relative_residuals)
        assert torch.equal(got.basis, want.basis)
    assert pair[0].converged and np.max(pair[0].shift_relative_residuals) <= 1e-10
    # the converged source's Green function equals the dense resolvent
    want = np.array([sink @ np.linalg.solve(zz * np.eye(P) - A, b1) for zz in z])
    assert _rel(pair[0].green, want) < 1e-8


def test_batched_arnoldi_rejects_plans_that_are_not_twins(torch_gpu):
    """The lockstep driver refuses bases whose plans were built from different arrays (two
    EOM problems of equal dimension), before it launches anything."""

    import ctypes

    import torch

    from c3lindblad_amd import _lib
    from c3lindblad_amd.eom import ComponentEOMProblem, component_eom_context, eom_ccsd_problem

    print("case batched arnoldi non-twin plans", flush=True)
    ops = []
    for seed in (31, 32):
        p, sp = _static(3, 1.0, seed)
        ep = ComponentEOMProblem.compile(eom_ccsd_problem(p), sp, 1)
        ops.append(component_eom_context(ep, np.zeros(p.num_amplitudes, complex)).device)
    n = ops[0].dim
    assert ops[1].dim == n and ops[0].plan.handle != ops[1].plan.handle
    m, ns, count = 4, 3, 2
    scratch_size, coef_size = _lib.krylov_workspace(True, n, m, ns)
    dev = "cuda"
    V = torch.zeros(count, m + 1, _lib.krylov_basis_ld(n), dtype=torch.complex128, device=dev)
    V[:, 0, 0] = 1.0
    H = torch.zeros(count, m, m + 1, dtype=torch.complex128, device=dev)
    zs = torch.full((ns,), 1j, dtype=torch.complex128, device=dev)
    sink = torch.ones(count, n, dtype=torch.complex128, device=dev)
    rot_c = torch.zeros(count, ns * m, dtype=torch.float64, device=dev)
    rot_s = torch.zeros(count, ns * m, dtype=torch.complex128, device=dev)
    pi = torch.zeros(count, ns * m, dtype=torch.complex128, device=dev)
    gamma = torch.zeros(count, ns, dtype=torch.complex128, device=dev)
    green = torch.zeros(count, ns, dtype=torch.complex128, device=dev)
    resid = torch.zeros(count, ns, dtype=torch.float64, device=dev)
    scratch = torch.zeros(count, scratch_size, dtype=torch.float64, device=dev)
    coef = torch.zeros(count, coef_size, dtype=torch.float64, device=dev)
    status = torch.zeros(count, 4, dtype=torch.float64, device=dev)
    plans = (ctypes.c_void_p * count)(*[op.plan.handle for op in ops])
    beta = (ctypes.c_double * count)(1.0, 1.0)
    m_out, broke, worst = (ctypes.c_int32 * count)(), (ctypes.c_int32 * count)(), (ctypes.c_double * count)()
    with pytest.raises(ValueError, match="twins"):
        _lib.call("c3l_arnoldi_batch", plans, count, _lib.dptr(ops[0].diag), _lib.dptr(V), _lib.dptr(H), n, m,
                  _lib.dptr(zs), ns, _lib.dptr(sink), beta, _lib.dptr(rot_c), _lib.dptr(rot_s), _lib.dptr(pi),
                  _lib.dptr(gamma), _lib.dptr(green), _lib.dptr(resid), _lib.dptr(scratch), _lib.dptr(coef),
                  _lib.dptr(status), 1e-10, 1, 1e-14, m_out, broke, worst, _lib.stream_ptr(torch))
    torch.cuda.synchronize()
    assert float(status.abs().sum()) == 0.0  # nothing was launched


def test_matrix_green_at_complex_shifts_matches_dense_oracle(torch_gpu):
    """Impurity Green matrix over two source sites at real-axis (omega + i eta) and
    imaginary-axis (i omega_n) shifts sharing one Krylov space per source and channel,
    against dense resolvents of the polynomial oracle's sigma matrix (1e-8 relative)."""

    from c3lindblad_amd import component_retarded_matrix_green_shifts
    from c3lindblad_amd.configs import readme_model
    from c3lindblad_amd.eom import (
        ComponentEOMProblem,
        ComponentSigmaOperator,
        component_eom_context,
        component_eom_sources,
        eom_ccsd_problem,
    )
    from c3lindblad_amd.lowering import factorized_cc_problem
    from c3lindblad_amd.solver import solve_component_ccsd
    from c3lindblad_amd.static import ComponentStaticProblem
    from oracle import polynomial as poly
    from oracle import solvers

    print("case matrix green at complex shifts", flush=True)
    p = factorized_cc_problem(readme_model())
    sp = ComponentStaticProblem.compile(p)
    sol = solve_component_ccsd(sp)
    ep = ComponentEOMProblem.compile(eom_ccsd_problem(p), sp, 1)
    ctx = component_eom_context(ep, sol.amplitudes)
    sources = [component_eom_sources(ctx, site=s) for s in (0, 1)]
    beta = 10.0
    z = np.concatenate([np.array([-1.0 + 0.2j, 0.3 + 0.5j, 2.5 + 0.05j]),
                        1j * np.pi / beta * (2 * np.arange(4) + 1)])
    op = ComponentSigmaOperator(ctx)
    out = component_retarded_matrix_green_shifts(op, sources, z, maximum_krylov_dimension=ep.dimension)
    masks = poly.coordinate_masks(ep.coordinates)
    smasks = poly.coordinate_masks(p.coordinates)
    A = solvers.dense_matrix(lambda v: poly.sigma(p, sol.amplitudes, masks, v, smasks), ep.dimension)
    eye = np.eye(ep.dimension)
    for name in ("addition", "removal"):
        want = np.empty((z.size, 2, 2), dtype=complex)
        for k, zk in enumerate(z):
            X = np.linalg.solve(zk * eye - A, np.stack([getattr(s, name) for s in sources], axis=1))
            want[k] = np.stack([s.sink for s in sources]) @ X
        assert _rel(getattr(out, name), want) < 1e-8, name
        assert np.max(getattr(out, f"{name}_relative_residuals")) < 1e-8
    assert out.sites == (0, 1) and out.retarded.shape == (z.size, 2, 2)
    assert out.action_calls == sum(p.dimension for p in out.addition_projections + out.removal_projections)
    with pytest.raises(ValueError):
        component_retarded_matrix_green_shifts(op, sources, np.array([0.5 + 0.0j]))  # Im z must be > 0


def test_time_domain_green_function_matches_dense_exponential(torch_gpu):
    """G(t) = -i sink . exp(-i t sigma)(b_add + b_rem) from one Arnoldi space against the dense
    matrix exponential of the polynomial oracle's sigma (1e-8 relative)."""

    import scipy.linalg

    from c3lindblad_amd import component_retarded_time
    from c3lindblad_amd.configs import readme_model
    from c3lindblad_amd.eom import (
        ComponentEOMProblem,
        ComponentSigmaOperator,
        component_eom_context,
        component_eom_sources,
        eom_ccsd_problem,
    )
    from c3lindblad_amd.lowering import factorized_cc_problem
    from c3lindblad_amd.solver import solve_component_ccsd
    from c3lindblad_amd.static import ComponentStaticProblem
    from oracle import polynomial as poly
    from oracle import solvers

    print("case time-domain green", flush=True)
    p = factorized_cc_problem(readme_model())
    sp = ComponentStaticProblem.compile(p)
    sol = solve_component_ccsd(sp)
    ep = ComponentEOMProblem.compile(eom_ccsd_problem(p), sp, 1)
    ctx = component_eom_context(ep, sol.amplitudes)
    sources = component_eom_sources(ctx)
    t = np.array([0.0, 0.3, 1.0, 2.5, 6.0])
    # a tolerance below rounding grows the space to all P = 39 directions, where the Krylov
    # exponential is exact; the reference tolerance (1e-11) is exercised on top
    got = component_retarded_time(ComponentSigmaOperator(ctx), sources, t, residual_tolerance=1e-15)
    default = component_retarded_time(ComponentSigmaOperator(ctx), sources, t)
    masks = poly.coordinate_masks(ep.coordinates)
    smasks = poly.coordinate_masks(p.coordinates)
    A = solvers.dense_matrix(lambda v: poly.sigma(p, sol.amplitudes, masks, v, smasks), ep.dimension)
    b = sources.addition + sources.removal
    want = np.array([-1j * sources.sink @ (scipy.linalg.expm(-1j * tk * A) @ b) for tk in t])
    assert _rel(got, want) < 1e-8
    assert _rel(default[:3], want[:3]) < 1e-6  # short times in the reference's converged space
    # G(0) = -i sink . b exactly (the basis spans b)
    assert abs(got[0] + 1j * (sources.sink @ b)) <= 1e-12 * max(1.0, abs(sources.sink @ b))

"""Reference resolvent API on the device: ``ArnoldiResolvent`` methods, the Galerkin stop rule,
the real-axis matrix drivers, the matrix time domain and the Matsubara Green function.

Every value is checked against dense linear algebra on the same sigma operator: the README
model's sigma matrix from the polynomial oracle (P = 39) and, where the Krylov space must stop
below the full dimension, a 6-site SIAM (P = 312) whose dense matrix is assembled column by
column from the device sigma (itself pinned against the port elsewhere).
"""

from __future__ import annotations

import numpy as np
import pytest
import scipy.linalg

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def readme(torch_gpu):
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

    print("fixture readme model", flush=True)
    p = factorized_cc_problem(readme_model())
    sp = ComponentStaticProblem.compile(p)
    sol = solve_component_ccsd(sp)
    ep = ComponentEOMProblem.compile(eom_ccsd_problem(p), sp, 1)
    ctx = component_eom_context(ep, sol.amplitudes)
    sources = [component_eom_sources(ctx, site=s) for s in (0, 1)]
    masks = poly.coordinate_masks(ep.coordinates)
    smasks = poly.coordinate_masks(p.coordinates)
    A = solvers.dense_matrix(lambda v: poly.sigma(p, sol.amplitudes, masks, v, smasks), ep.dimension)
    return dict(ctx=ctx, sources=sources, A=A, P=ep.dimension, op=lambda: ComponentSigmaOperator(ctx))


@pytest.fixture(scope="module")
def siam6(torch_gpu):
    import torch

    from c3lindblad_amd.configs import siam_model
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

    print("fixture siam n=6", flush=True)
    p = factorized_cc_problem(siam_model(5, 4.0, seed=7))
    sp = ComponentStaticProblem.compile(p)
    sol = solve_component_ccsd(sp)
    assert sol.converged
    ep = ComponentEOMProblem.compile(eom_ccsd_problem(p), sp, 1)
    ctx = component_eom_context(ep, sol.amplitudes)
    op = ComponentSigmaOperator(ctx)
    P = ep.dimension
    eye = torch.eye(P, dtype=torch.complex128, device="cuda")
    A = torch.empty(P, P, dtype=torch.complex128, device="cuda")
    for k in range(P):
        op.matvec_device(eye[k].contiguous(), A[k])
    return dict(ctx=ctx, sources=[component_eom_sources(ctx, site=0)], A=A.cpu().numpy().T, P=P,
                op=lambda: ComponentSigmaOperator(ctx))


def test_arnoldi_resolvent_has_the_reference_methods(readme):
    """coefficients / residual_norms / relative_residuals / solutions / transfer / time_coefficients /
    to_dict of one Galerkin-stopped space against dense solves; the residual the driver stops on
    equals the Schur-form residual and the exact residual of the reconstructed solutions."""

    from c3lindblad_amd import arnoldi_resolvent

    print("case arnoldi resolvent reference api", flush=True)
    A, P = readme["A"], readme["P"]
    rng = np.random.default_rng(3)
    b = rng.normal(size=P) + 1j * rng.normal(size=P)
    sink = rng.normal(size=P) + 1j * rng.normal(size=P)
    z = np.concatenate([np.linspace(-3.0, 3.0, 7) + 0.2j, 1j * np.pi / 8.0 * (2 * np.arange(3) + 1)])
    proj = arnoldi_resolvent(readme["op"](), b, z, maximum_dimension=P, residual_tolerance=1e-10, minimum_dimension=4)
    m = proj.dimension
    assert proj.converged and proj.criterion == "galerkin"
    rel = proj.relative_residuals(z)
    assert np.max(rel) <= 1e-10
    assert np.allclose(rel, proj.shift_relative_residuals, rtol=1e-5, atol=1e-14)
    if m > 4:  # the driver stops at the first passing dimension
        assert np.max(proj.relative_residuals(z, dimension=m - 1)) > 1e-10
    X = proj.solutions(z)
    eye = np.eye(P)
    want = np.stack([np.linalg.solve(zk * eye - A, b) for zk in z])
    assert _rel(X, want) < 1e-8
    true = np.array([np.linalg.norm(b - (zk * eye - A) @ x) for zk, x in zip(z, X)]) / np.linalg.norm(b)
    assert np.allclose(true, rel, rtol=1e-4, atol=1e-13)
    assert _rel(proj.transfer(z, sink), want @ sink) < 1e-8
    # a saved prefix equals the solve on the leading block of H
    k = max(1, m - 3)
    y = proj.coefficients(z[:2], dimension=k)
    rhs = np.zeros(k, dtype=complex)
    rhs[0] = proj.source_norm
    for zk, yk in zip(z[:2], y):
        assert np.allclose(yk, np.linalg.solve(zk * np.eye(k) - proj.hessenberg[:k, :k], rhs), rtol=1e-10, atol=1e-13)
    assert proj.orthogonality_defect < 1e-13
    record = proj.to_dict()
    assert set(record) == {
        "krylov_dimension", "full_vector_dimension", "source_norm", "terminal_coupling_absolute",
        "orthogonality_defect", "action_wall_seconds", "action_wall_seconds_total", "construction_wall_seconds",
        "peak_rss_kib", "breakdown", "check_dimensions", "check_maximum_relative_residuals",
        "maximum_relative_residual", "residual_tolerance", "converged", "method",
    }
    assert record["krylov_dimension"] == m and record["full_vector_dimension"] == P
    # the whole space: the Krylov exponential is exact
    full = arnoldi_resolvent(readme["op"](), b, z[:1], maximum_dimension=P, residual_tolerance=1e-300,
                             minimum_dimension=P)
    t = np.array([0.0, 0.7, 2.0])
    got = full.time_coefficients(t) @ full.basis[: full.dimension].cpu().numpy()
    want_t = np.stack([scipy.linalg.expm(-1j * tk * A) @ b for tk in t])
    assert full.dimension == P or full.breakdown
    assert _rel(got, want_t) < 1e-9


def test_galerkin_stop_below_the_full_dimension(siam6):
    """A Krylov cap below P (120 of 312): every shift's Galerkin residual meets the tolerance at the
    returned dimension (and not one column earlier), and every Green-matrix value agrees with the
    dense resolvent within what that residual allows."""

    from c3lindblad_amd import component_retarded_matrix_green_shifts

    print("case galerkin stop below full dimension", flush=True)
    A, P, sources = siam6["A"], siam6["P"], siam6["sources"]
    z = np.concatenate([np.linspace(-4.0, 4.0, 41) + 0.05j, 1j * np.pi / 20.0 * (2 * np.arange(8) + 1)])
    out = component_retarded_matrix_green_shifts(siam6["op"](), sources, z, maximum_krylov_dimension=200,
                                                 residual_tolerance=1e-9)
    eye = np.eye(P)
    for name in ("addition", "removal"):
        proj = getattr(out, f"{name}_projections")[0]
        m = proj.dimension
        print(name, "dimension", m, "of", P, flush=True)
        assert m < P and proj.converged
        assert np.max(getattr(out, f"{name}_relative_residuals")) <= 1e-9
        assert np.max(proj.relative_residuals(z, dimension=m - 1)) > 1e-9
        b = getattr(sources[0], name)
        want = np.array([sources[0].sink @ np.linalg.solve(zk * eye - A, b) for zk in z])
        assert _rel(getattr(out, name)[:, 0, 0], want) < 1e-6


def test_real_axis_matrix_drivers_match_dense_resolvents(readme):
    """component_retarded_matrix_green (per-frequency acceptance) and
    component_retarded_matrix_green_arnoldi on omega + i eta, axes (omega, sink, source)."""

    from c3lindblad_amd import component_retarded_matrix_green, component_retarded_matrix_green_arnoldi

    print("case real-axis matrix drivers", flush=True)
    A, P, sources = readme["A"], readme["P"], readme["sources"]
    omega = np.linspace(-4.0, 4.0, 17)
    eta = 0.05
    eye = np.eye(P)
    sinks = np.stack([s.sink for s in sources])
    direct = component_retarded_matrix_green(readme["op"](), sources, omega, eta)
    shared = component_retarded_matrix_green_arnoldi(readme["op"](), sources, omega, eta)
    for name in ("addition", "removal"):
        B = np.stack([getattr(s, name) for s in sources], axis=1)
        want = np.stack([sinks @ np.linalg.solve((w + 1j * eta) * eye - A, B) for w in omega])
        assert _rel(getattr(direct, name), want) < 1e-8, name
        assert _rel(getattr(shared, name), want) < 1e-8, name
        records = getattr(direct, f"{name}_solutions")
        assert len(records) == 2 and all(len(r) == omega.size for r in records)
        assert all(r.converged and r.true_relative_residual <= 1e-9 for rec in records for r in rec)
    assert np.array_equal(direct.retarded, direct.addition + direct.removal)
    assert shared.component_action_calls == sum(p.dimension for p in shared.addition_projections
                                                + shared.removal_projections)
    with pytest.raises(ValueError):
        component_retarded_matrix_green(readme["op"](), sources, omega, 0.0)  # eta must be positive


def test_matrix_time_domain_matches_dense_exponential(readme):
    """component_retarded_matrix_time on the whole space: G_ab(t) = -i sink_a exp(-i t sigma) b_b,
    the channel split adds up to the combined column, and negative times are rejected."""

    from c3lindblad_amd import component_retarded_matrix_time

    print("case matrix time domain", flush=True)
    A, P, sources = readme["A"], readme["P"], readme["sources"]
    t = np.array([0.0, 0.5, 1.5, 4.0])
    out = component_retarded_matrix_time(readme["op"](), sources, t, np.array([1j]), residual_tolerance=1e-300,
                                         maximum_krylov_dimension=P)
    sinks = np.stack([s.sink for s in sources])
    B = np.stack([s.addition + s.removal for s in sources], axis=1)
    want = np.stack([-1j * sinks @ (scipy.linalg.expm(-1j * tk * A) @ B) for tk in t])
    assert _rel(out.retarded, want) < 1e-8
    assert out.channel_sum_defect < 1e-10 * np.max(np.abs(want))
    assert out.retarded.shape == (t.size, 2, 2) and out.sites == (0, 1)
    with pytest.raises(ValueError):
        component_retarded_matrix_time(readme["op"](), sources, np.array([-1.0]), np.array([1j]))


def test_matsubara_green_matches_dense_resolvents(readme):
    """G(i omega_n) on a symmetric grid: positive points from one Krylov space per source and
    channel, negative points the conjugate transpose."""

    from c3lindblad_amd import component_matsubara_green, matsubara_grid

    print("case matsubara green", flush=True)
    A, P, sources = readme["A"], readme["P"], readme["sources"]
    grid = matsubara_grid(10.0, n_iw=4)
    out = component_matsubara_green(readme["op"](), sources, grid)
    eye = np.eye(P)
    sinks = np.stack([s.sink for s in sources])
    B = np.stack([s.addition + s.removal for s in sources], axis=1)
    upper = np.stack([sinks @ np.linalg.solve(1j * abs(w) * eye - A, B) for w in grid.omega_n])
    want = np.where((grid.indices < 0)[:, None, None], upper.conj().swapaxes(1, 2), upper)
    assert _rel(out.green, want) < 1e-8
    assert out.max_relative_residual <= 1e-10


def test_checkpointed_arnoldi_resumes_exactly(siam6, tmp_path):
    """An interrupted (maximum_new_steps) and resumed disk-backed run equals the uninterrupted run bit
    for bit (``tests/green/test_checkpointed_arnoldi.py:67-135`` checks in-memory against
    checkpointed the same way); its transfer matches dense resolvents; a checkpoint refuses other
    inputs."""

    from c3lindblad_amd import checkpointed_arnoldi_resolvent, eom_checkpoint_identity

    print("case checkpointed arnoldi n=6", flush=True)
    A, P, sources = siam6["A"], siam6["P"], siam6["sources"]
    b = sources[0].addition
    sinks = np.stack([sources[0].sink, np.roll(sources[0].sink, 3)])
    z = np.linspace(-3.0, 3.0, 13) + 0.1j
    ident = eom_checkpoint_identity(siam6["ctx"])
    kw = dict(maximum_dimension=160, checkpoint_dimensions=(20, 40, 60, 80, 100, 120, 140, 160),
              residual_tolerance=1e-10, provenance=ident)
    whole = checkpointed_arnoldi_resolvent(siam6["op"](), b, sinks, z, tmp_path / "whole", **kw)
    part = checkpointed_arnoldi_resolvent(siam6["op"](), b, sinks, z, tmp_path / "split", maximum_new_steps=33, **kw)
    assert part.projection.dimension == 33 and not part.projection.converged
    done = checkpointed_arnoldi_resolvent(siam6["op"](), b, sinks, z, tmp_path / "split", **kw)
    m = whole.projection.dimension
    print("dimension", m, whole.projection.check_dimensions, flush=True)
    assert whole.projection.converged and done.projection.dimension == m
    assert np.array_equal(done.projection.hessenberg, whole.projection.hessenberg)
    assert np.array_equal(done.sink_projections, whole.sink_projections)
    assert np.array_equal(done.transfer(z), whole.transfer(z))
    mirror = np.load(tmp_path / "split" / "basis.npy", mmap_mode="r")
    assert np.array_equal(mirror[:, : m + 1], whole.projection.basis.cpu().numpy().T)
    eye = np.eye(P)
    want = np.stack([sinks @ np.linalg.solve(zk * eye - A, b) for zk in z])
    assert _rel(whole.transfer(z), want) < 1e-8
    with pytest.raises(ValueError, match="do not match"):
        checkpointed_arnoldi_resolvent(siam6["op"](), b, sinks, z, tmp_path / "split",
                                       **{**kw, "residual_tolerance": 1e-9})
    with pytest.raises(ValueError, match="zero new steps"):
        checkpointed_arnoldi_resolvent(siam6["op"](), b, sinks, z, tmp_path / "fresh", maximum_new_steps=0, **kw)

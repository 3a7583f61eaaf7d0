"""Pins of the CPU oracle (no GPU).

* polynomial oracle residual / sigma == sparse Fock-space brute force of
  ``exp(-T) K exp(T)`` (every sign from literal Jordan--Wigner products);
* polynomial JVP == central finite difference of the polynomial residual
  (``tests/cc/test_charge_s13.py:125-135`` uses the same check);
* compiled plans executed by the NumPy port == polynomial oracle (pins the
  term compiler, the plan layout and the padded-tensor conventions);
* the residual is tilde-even (``t_partner = η conj(t)``), so packing loses
  nothing (``src/c3l/cc/factorized.py:16-19``).
"""

from __future__ import annotations

import numpy as np
import pytest

from tests.conftest import random_model


def _setup(n, U, seed):
    """``n = "kanamori3"`` / ``"kanamori4"``: two Kanamori orbitals with exchange, pair hopping and a
    random 8-fold symmetric V_pqrs (a multi-site general ``two_body_interaction``) plus one or
    two pseudomodes; otherwise a random single-impurity model."""

    from c3lindblad_amd.configs import multiorbital_model
    from c3lindblad_amd.lowering import factorized_cc_problem
    from c3lindblad_amd.static import ComponentStaticProblem

    if n == "kanamori3":
        model = multiorbital_model(2, (1, 0), seed=3, U=2.0, J=0.3, perturbation=0.1, impurity_hopping=0.3)
    elif n == "kanamori4":
        model = multiorbital_model(2, (1, 1), seed=1, U=2.0, J=0.3, perturbation=0.1, impurity_hopping=0.2)
    else:
        model = random_model(n, U, seed)
    p = factorized_cc_problem(model)
    return p, ComponentStaticProblem.compile(p)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("n,U,seed", [(1, 1.3, 1), (2, 0.9, 2), (3, 1.2, 3), ("kanamori3", 0, 3)])
def test_polynomial_residual_equals_fock_brute_force(n, U, seed):
    from oracle import polynomial as poly
    from oracle.fock import PHFockSpace

    p, _ = _setup(n, U, seed)
    rng = np.random.default_rng(seed)
    t = p.tilde.unpack(0.1 * rng.normal(size=p.num_variables))
    R, _ = poly.residual(p, t)
    assert _rel(PHFockSpace(p).residual(p.coordinates, t), R) < 1e-13


@pytest.mark.parametrize("n,charge", [(1, 1), (2, 1), (2, -1), (3, 1), (3, -1), ("kanamori3", 1), ("kanamori3", -1)])
def test_polynomial_sigma_equals_fock_brute_force(n, charge):
    from c3lindblad_amd.lowering import response_coordinates
    from oracle import polynomial as poly
    from oracle.fock import PHFockSpace

    p, _ = _setup(n, 0.8, 30 + (n if isinstance(n, int) else 3))
    rng = np.random.default_rng(n if isinstance(n, int) else 5)
    t = p.tilde.unpack(0.1 * rng.normal(size=p.num_variables))
    rc = response_coordinates(p.model.n_sites, charge)
    x = rng.normal(size=rc.dimension) + 1j * rng.normal(size=rc.dimension)
    want = PHFockSpace(p).sigma(p.coordinates, t, rc, x)
    assert _rel(poly.sigma(p, t, poly.coordinate_masks(rc), x), want) < 1e-13


@pytest.mark.parametrize("n,charge", [(2, 1), (2, -1), (3, 1), (3, -1)])
def test_polynomial_s135_sigma_equals_fock_brute_force(n, charge):
    """The full S1 + S3 + S5 sigma (rank-five shell, ``cc/s135_full.py:101-361``) of the
    polynomial oracle against the Fock-space brute force of ``[K̄, X]|0>`` projected on the S135
    strings (n = 2: 22 strings per charge, n = 3: 150)."""

    from c3lindblad_amd.lowering import response_coordinates
    from oracle import polynomial as poly
    from oracle.fock import PHFockSpace

    p, _ = _setup(n, 0.8, 50 + n)
    rng = np.random.default_rng(n + 11)
    t = p.tilde.unpack(0.1 * rng.normal(size=p.num_variables))
    rc = response_coordinates(p.model.n_sites, charge, "S135")
    assert any(sum(b.key) == 5 for b in rc.blocks)
    x = rng.normal(size=rc.dimension) + 1j * rng.normal(size=rc.dimension)
    want = PHFockSpace(p).sigma(p.coordinates, t, rc, x)
    assert _rel(poly.sigma(p, t, poly.coordinate_masks(rc), x, max_rank=5), want) < 1e-13


@pytest.mark.parametrize("n,charge", [(2, 1), (3, -1), (4, 1)])
def test_compiled_s135_sigma_equals_polynomial_oracle(n, charge):
    """The S135 plan (rank-five input, output and creation tensors) executed by the NumPy port
    against the polynomial oracle through rank five."""

    from c3lindblad_amd.eom import ComponentEOMProblem, eom_ccsd_problem
    from oracle import polynomial as poly
    from oracle.port import PortPlan

    p, sp = _setup(n, 1.7, 70 + n)
    rng = np.random.default_rng(n + 13)
    t = p.tilde.unpack(0.05 * rng.normal(size=p.num_variables))
    ep = ComponentEOMProblem.compile(eom_ccsd_problem(p, space="S135"), sp, charge)
    x = rng.normal(size=ep.dimension) + 1j * rng.normal(size=ep.dimension)
    port = PortPlan(ep.lowered)
    port.load(0, t)
    port.load(2, x)
    got = port.apply(3, np.zeros(ep.dimension, complex), ep.diagonal, x)
    assert _rel(got, poly.sigma(p, t, poly.coordinate_masks(ep.coordinates), x, max_rank=5)) < 1e-13


def test_polynomial_jvp_equals_finite_difference():
    from oracle import polynomial as poly

    p, _ = _setup(2, 1.5, 8)
    rng = np.random.default_rng(3)
    t = p.tilde.unpack(0.05 * rng.normal(size=p.num_variables))
    d = p.tilde.unpack(rng.normal(size=p.num_variables))
    h = 1e-5
    fd = (poly.residual(p, t + h * d)[0] - poly.residual(p, t - h * d)[0]) / (2 * h)
    assert np.max(np.abs(poly.jvp(p, t, d) - fd)) < 1e-9


@pytest.mark.parametrize("n,U,seed", [(1, 1.3, 1), (2, 0.9, 2), (3, 1.2, 3), (4, 2.5, 4), ("kanamori3", 0, 3),
                                     ("kanamori4", 0, 4)])
def test_compiled_plan_equals_polynomial_oracle(n, U, seed):
    from oracle import polynomial as poly
    from oracle.port import PortPlan

    p, sp = _setup(n, U, seed)
    rng = np.random.default_rng(seed + 100)
    t = p.tilde.unpack(0.05 * rng.normal(size=p.num_variables))
    dt = p.tilde.unpack(rng.normal(size=p.num_variables))
    R, ref = poly.residual(p, t)
    port = PortPlan(sp.lowered)
    port.load(0, t)
    got = port.apply(0, np.zeros(p.num_amplitudes, complex), sp.diagonal, t)
    assert _rel(got, R) < 1e-13
    got_ref = port.apply(1, np.zeros(1, complex))
    assert abs(got_ref[0] - ref) < 1e-13 * max(1.0, abs(ref))
    port.load(1, dt)
    assert _rel(port.apply(2, np.zeros(p.num_amplitudes, complex), sp.diagonal, dt), poly.jvp(p, t, dt)) < 1e-13
    assert p.tilde.defect(R) < 1e-13 * max(1.0, np.max(np.abs(R)))


@pytest.mark.parametrize("n,U,seed", [(2, 0.9, 2), (3, 1.2, 3), ("kanamori3", 0, 3)])
def test_simpson_residual_equals_polynomial_oracle(n, U, seed):
    """The residual is a polynomial of degree four in the amplitudes, so ``f(s) = J(s t) t`` is a
    cubic and Simpson's rule integrates it exactly: ``R(t) = R(0) + (f(0) + 4 f(1/2) + f(1))/6``
    with the oracle's own residual and JVP.  The compiled ``residual_low`` program (degree-zero
    terms, one sixth of the linear ones, a sixth of the diagonal) equals ``R(0) + f(0)/6`` of the
    oracle, and the three compiled programs assembled by the rule give the oracle's residual."""

    from oracle import polynomial as poly
    from oracle.port import PortPlan
    from c3lindblad_amd.plan_arrays import PROGRAMS

    p, sp = _setup(n, U, seed)
    assert sp.residual_degree == 4
    rng = np.random.default_rng(seed + 300)
    t = p.tilde.unpack(0.05 * rng.normal(size=p.num_variables))
    R, _ = poly.residual(p, t)
    zero = np.zeros_like(t)
    f = [poly.jvp(p, s * t, t) for s in (0.0, 0.5, 1.0)]
    assert _rel(poly.residual(p, zero)[0] + (f[0] + 4 * f[1] + f[2]) / 6, R) < 1e-13
    # the trapezoid rule is not exact here: the identity above is not satisfied by accident
    assert _rel(poly.residual(p, zero)[0] + (f[0] + f[2]) / 2, R) > 1e-6
    port = PortPlan(sp.lowered)
    port.load(0, t)
    low = port.apply(PROGRAMS["residual_low"], np.zeros(p.num_amplitudes, complex), sp.diagonal / 6, t)
    assert _rel(low, poly.residual(p, zero)[0] + f[0] / 6) < 1e-13
    port.load(1, t)
    f1 = port.apply(PROGRAMS["jvp"], np.zeros(p.num_amplitudes, complex), sp.diagonal, t)
    port.load(0, 0.5 * t)
    fh = port.apply(PROGRAMS["jvp"], np.zeros(p.num_amplitudes, complex), sp.diagonal, t)
    assert _rel(low + (4 * fh + f1) / 6, R) < 1e-13


@pytest.mark.parametrize("n,charge", [(2, 1), (3, -1), (4, 1), ("kanamori3", 1), ("kanamori4", -1)])
def test_compiled_sigma_and_sources_equal_polynomial_oracle(n, charge):
    from c3lindblad_amd.eom import ComponentEOMProblem, _sink, channel_operators, eom_ccsd_problem
    from c3lindblad_amd.lowering import physical_mode_operator
    from oracle import polynomial as poly
    from oracle.port import PortPlan

    p, sp = _setup(n, 1.7, 40 + (n if isinstance(n, int) else 0))
    n = p.model.n_sites
    rng = np.random.default_rng(n + 7)
    t = p.tilde.unpack(0.05 * rng.normal(size=p.num_variables))
    ep = ComponentEOMProblem.compile(eom_ccsd_problem(p), sp, charge)
    masks = poly.coordinate_masks(ep.coordinates)
    x = rng.normal(size=ep.dimension) + 1j * rng.normal(size=ep.dimension)
    port = PortPlan(ep.lowered)
    port.load(0, t)
    port.load(2, x)
    got = port.apply(3, np.zeros(ep.dimension, complex), ep.diagonal, x)
    assert _rel(got, poly.sigma(p, t, masks, x)) < 1e-13
    for channel in ("c", "c_dagger", "tilde_c", "tilde_c_dagger"):
        comps = physical_mode_operator(p.reference, 0, 0, channel).components
        assert np.array_equal(_sink(ep, comps), poly.sink(masks, comps))
    for site in range(n):
        ops = channel_operators(p.reference, site, 0, charge)
        port = PortPlan(ep.source_plan(site))
        port.load(0, t)
        for name in ("addition", "removal"):
            op, coefficient = ops[name]
            want = coefficient * poly.source(p, t, masks, op.components)
            got = port.apply(6 if name == "addition" else 7, np.zeros(ep.dimension, complex))
            assert np.max(np.abs(got - want)) < 1e-14 * max(1.0, np.max(np.abs(want)))
        assert np.array_equal(ep.sink(site), poly.sink(masks, ops["sink"][0].components))


def test_empty_and_degenerate_inputs_are_rejected():
    from oracle import polynomial as poly

    p, sp = _setup(2, 1.0, 2)
    t = np.zeros(p.num_amplitudes, complex)
    R, ref = poly.residual(p, t)
    # the bare residual at T = 0 is the vacuum projection of K; it must be finite
    assert np.all(np.isfinite(R)) and np.isfinite(ref)
    with pytest.raises(ValueError):
        p.tilde.unpack(np.zeros(p.num_variables + 1))
    with pytest.raises(ValueError):
        p.tilde.pack(np.full(p.num_amplitudes, np.nan))


def _mode_ops(p):
    from c3lindblad_amd.lowering import physical_mode_operator

    ref = p.reference
    N = ref.model.n_spin_orbitals
    ann = [physical_mode_operator(ref, *ref.model.orbital_site_spin(o), "c").components for o in range(N)]
    cre = [physical_mode_operator(ref, *ref.model.orbital_site_spin(o), "c_dagger").components for o in range(N)]
    return ann, cre


@pytest.mark.parametrize("n,U,seed", [(1, 1.3, 1), (2, 0.9, 2)])
def test_low_rank_prefix_expectations_equal_complete_exponential(n, U, seed):
    """Products of <= 4 linear operators see only ranks 0/2/4 of exp(T)|0>: the
    prefix 1 + T + T1∧T1/2 must equal the complete exponential for them."""

    from oracle import polynomial as poly
    from oracle.fock import PHFockSpace

    p, _ = _setup(n, U, seed)
    rng = np.random.default_rng(seed + 10)
    t = p.tilde.unpack(0.2 * rng.normal(size=p.num_variables))
    masks = poly.coordinate_masks(p.coordinates)
    prefix = poly.low_rank_prefix(masks, t)
    space = PHFockSpace(p)
    ann, cre = _mode_ops(p)
    N = len(ann)
    products = [(cre[q], ann[r]) for q in range(N) for r in range(N)]
    products += [(cre[a], ann[b], cre[c], ann[d]) for a, b, c, d in rng.integers(0, N, size=(12, 4))]
    got = np.array([poly.expectation(prefix, f) for f in products])
    want = np.array([space.expectation(p.coordinates, t, f) for f in products])
    assert np.max(np.abs(got - want)) < 1e-13 * max(1.0, np.max(np.abs(want)))
    # the T1∧T1/2 term matters: dropping it changes the quartic values
    bare = {0: 1.0 + 0.0j, **{m: complex(v) for m, v in zip(masks, t)}}
    assert max(abs(poly.expectation(bare, f) - w) for f, w in zip(products, want)) > 1e-4


@pytest.mark.parametrize("n,U,seed", [(1, 1.3, 1), (2, 0.9, 2), (3, 1.2, 3)])
def test_compiled_observables_equal_polynomial_oracle(n, U, seed):
    from c3lindblad_amd.lowering import FactorizedModeProduct, physical_mode_operator
    from c3lindblad_amd.observables import ComponentObservableProblem, covariance_problem
    from oracle import polynomial as poly
    from oracle.port import PortPlan

    p, sp = _setup(n, U, seed)
    rng = np.random.default_rng(seed + 20)
    t = p.tilde.unpack(0.1 * rng.normal(size=p.num_variables))
    cp = covariance_problem(sp)
    port = PortPlan(cp.lowered)
    port.load(0, t)
    out = np.zeros(len(cp.products), dtype=np.complex128)
    port.apply(5, out)
    want = poly.covariance(p, t)
    assert _rel(out.reshape(want.shape), want) < 1e-13
    ref = p.reference
    N = ref.model.n_spin_orbitals
    quartets = rng.integers(0, N, size=(10, 4))
    ops = {}
    for o in range(N):
        s = ref.model.orbital_site_spin(o)
        ops[o, 0] = physical_mode_operator(ref, *s, "c")
        ops[o, 1] = physical_mode_operator(ref, *s, "c_dagger")
    products = [FactorizedModeProduct("q", 1.0, (ops[a, 1], ops[b, 0], ops[c, 1], ops[d, 0])) for a, b, c, d in quartets]
    op = ComponentObservableProblem.compile(sp, products)
    port = PortPlan(op.lowered)
    port.load(0, t)
    out = np.zeros(len(products), dtype=np.complex128)
    port.apply(5, out)
    prefix = poly.low_rank_prefix(poly.coordinate_masks(p.coordinates), t)
    want = np.array([poly.expectation(prefix, tuple(f.components for f in pr.factors)) for pr in products])
    assert np.max(np.abs(out - want)) < 1e-13 * max(1.0, np.max(np.abs(want)))


@pytest.mark.parametrize("n,U,seed", [(2, 0.0, 4), (3, 1.1, 6)])
def test_covariance_at_zero_amplitudes_is_the_reference_covariance(n, U, seed):
    from oracle import polynomial as poly

    p, _ = _setup(n, U, seed)
    got = poly.covariance(p, np.zeros(p.num_amplitudes, dtype=np.complex128))
    assert np.max(np.abs(got - p.reference.gamma)) < 1e-13


def test_port_sampled_program_equals_full_program():
    """``PortPlan.apply_sampled`` (the spot check used where the full port residual is too slow)
    returns exactly the full program's values at its coordinates."""

    from oracle.port import PortPlan

    p, sp = _setup(4, 2.5, 4)
    rng = np.random.default_rng(9)
    t = p.tilde.unpack(0.05 * rng.normal(size=p.num_variables))
    port = PortPlan(sp.lowered)
    port.load(0, t, creation=False)
    full = port.apply(0, np.zeros(p.num_amplitudes, complex))
    dest, values = port.apply_sampled(0, 7, seed=1)
    assert dest.size == len({int(d) for d in dest}) and dest.size < p.num_amplitudes
    assert np.array_equal(values, full[dest])


def test_union_plan_equals_its_parts():
    """A union of three plans of equal n (different models, one with U = 0 and so fewer products)
    executed by the port gives every part's residual, JVP and sigma side by side."""

    from c3lindblad_amd.eom import ComponentEOMProblem, eom_ccsd_problem
    from c3lindblad_amd.plan_arrays import PROGRAMS, union_plan
    from oracle.port import PortPlan

    parts = [_setup(3, U, seed) for U, seed in ((1.5, 11), (0.0, 12), (2.2, 13))]
    D = parts[0][0].num_amplitudes
    rng = np.random.default_rng(4)
    ts = [p.tilde.unpack(0.05 * rng.normal(size=D)) for p, _ in parts]
    dts = [p.tilde.unpack(rng.normal(size=D)) for p, _ in parts]
    u = union_plan([sp.lowered for _, sp in parts], {0: D, 1: D}, {0: D, 1: 1, 2: D, 8: D})
    port = PortPlan(u)
    port.load(0, np.concatenate(ts))
    port.load(1, np.concatenate(dts))
    R = port.apply(0, np.zeros(3 * D, complex))
    ref = port.apply(1, np.zeros(3, complex))
    J = port.apply(2, np.zeros(3 * D, complex))
    eoms = [ComponentEOMProblem.compile(eom_ccsd_problem(p), sp, -1) for p, sp in parts]
    P = eoms[0].dimension
    xs = [rng.normal(size=P) + 1j * rng.normal(size=P) for _ in parts]
    ue = union_plan([e.lowered for e in eoms], {0: D, 2: P}, {PROGRAMS["sigma"]: P})
    eport = PortPlan(ue)
    eport.load(0, np.concatenate(ts))
    eport.load(2, np.concatenate(xs))
    S = eport.apply(3, np.zeros(3 * P, complex))
    for i, ((p, sp), e) in enumerate(zip(parts, eoms)):
        one = PortPlan(sp.lowered)
        one.load(0, ts[i])
        one.load(1, dts[i])
        assert np.array_equal(R[i * D:(i + 1) * D], one.apply(0, np.zeros(D, complex)))
        assert ref[i] == one.apply(1, np.zeros(1, complex))[0]
        assert np.array_equal(J[i * D:(i + 1) * D], one.apply(2, np.zeros(D, complex)))
        single = PortPlan(e.lowered)
        single.load(0, ts[i])
        single.load(2, xs[i])
        assert np.array_equal(S[i * P:(i + 1) * P], single.apply(3, np.zeros(P, complex)))
    assert len(parts[1][1].lowered.blocks[0][4]) < len(parts[0][1].lowered.blocks[0][4])  # U = 0: fewer entries


def test_config1_matches_the_reference_golden_vectors():
    """The README quick-start model solved by the reference's own NumPy path
    (``tests/golden/make_config1_reference_vectors.py``, run in the build container) against the
    oracle: the same amplitude masks in the same order, the same rapidities, the same converged
    amplitudes in the reference's gauge (full-HF reference and normal-mode canonicalisation,
    which no brute-force check fixes), the same charge +1 and -1 G^R on 81 frequencies and the
    same occupations."""

    from pathlib import Path

    from c3lindblad_amd.configs import readme_model
    from c3lindblad_amd.eom import ComponentEOMProblem, channel_operators, eom_ccsd_problem
    from c3lindblad_amd.lowering import factorized_cc_problem
    from c3lindblad_amd.static import ComponentStaticProblem
    from oracle import polynomial as poly
    from oracle import solvers

    g = np.load(Path(__file__).parent / "golden" / "config1_reference_vectors.npz")
    model = readme_model()
    p = factorized_cc_problem(model)
    masks = poly.coordinate_masks(p.coordinates)
    assert masks == [int(m) for m in g["masks"]]
    assert np.max(np.abs(np.asarray(p.rapidities) - g["rapidities"])) < 1e-12
    _, t, norm, _, _ = solvers.solve_ccsd(p.tilde, p.rapidity_sums(), lambda a: poly.residual(p, a)[0],
                                          lambda a: (lambda d: poly.jvp(p, a, d)))
    assert norm <= 1e-10 and float(g["packed_residual_norm"]) <= 1e-10
    assert np.max(np.abs(t - g["amplitudes"])) < 1e-12 * np.max(np.abs(g["amplitudes"]))
    sp = ComponentStaticProblem.compile(p)
    ep = ComponentEOMProblem.compile(eom_ccsd_problem(p), sp, 1)
    rmasks = poly.coordinate_masks(ep.coordinates)
    ops = channel_operators(p.reference, model.impurity_site, 0, 1)
    add = ops["addition"][1] * poly.source(p, t, rmasks, ops["addition"][0].components)
    rem = ops["removal"][1] * poly.source(p, t, rmasks, ops["removal"][0].components)
    sink = poly.sink(rmasks, ops["sink"][0].components)
    A = solvers.dense_matrix(lambda v: poly.sigma(p, t, rmasks, v, masks), ep.dimension)
    ga, gr = solvers.retarded_green_dense(A, add, rem, sink, g["omega"], 0.0)
    assert _rel(ga + gr, g["retarded"]) < 1e-11
    # the charge -1 block and the occupations of the same root
    em = ComponentEOMProblem.compile(eom_ccsd_problem(p), sp, -1)
    mmasks = poly.coordinate_masks(em.coordinates)
    ops = channel_operators(p.reference, model.impurity_site, 0, -1)
    add = ops["addition"][1] * poly.source(p, t, mmasks, ops["addition"][0].components)
    rem = ops["removal"][1] * poly.source(p, t, mmasks, ops["removal"][0].components)
    sink = poly.sink(mmasks, ops["sink"][0].components)
    A = solvers.dense_matrix(lambda v: poly.sigma(p, t, mmasks, v, masks), em.dimension)
    ga, gr = solvers.retarded_green_dense(A, add, rem, sink, g["omega"], 0.0)
    assert _rel(ga + gr, g["retarded_minus"]) < 1e-11
    assert _rel(poly.covariance(p, t), g["covariance"]) < 1e-13


@pytest.mark.parametrize("name", ["siam6", "multiorbital6"])
def test_reference_roots_solve_the_compiled_equations(name):
    """Two n = 6 models of the benchmark families (a SIAM with five pseudomodes at U = 4, and two
    Kanamori orbitals with V_pqrs and two pseudomodes each), solved by the reference's NumPy path
    from their model records (``tests/golden/make_config1_reference_vectors.py``): the same
    amplitude masks in the same order, the same rapidities, and the reference's converged
    amplitudes make this build's compiled residual vanish (packed norm <= 1e-12), so the two
    solve the same equations in the same coordinates and gauge."""

    from pathlib import Path

    from c3lindblad_amd.configs import multiorbital_model, siam_model
    from c3lindblad_amd.lowering import factorized_cc_problem
    from c3lindblad_amd.static import ComponentStaticProblem
    from oracle import polynomial as poly
    from oracle.port import PortPlan

    model = siam_model(5, 4.0, seed=3) if name == "siam6" else multiorbital_model(2, 2, seed=3)
    g = np.load(Path(__file__).parent / "golden" / f"{name}_reference_vectors.npz")
    p = factorized_cc_problem(model)
    assert poly.coordinate_masks(p.coordinates) == [int(m) for m in g["masks"]]
    assert np.max(np.abs(np.asarray(p.rapidities) - g["rapidities"])) < 1e-12
    sp = ComponentStaticProblem.compile(p)
    port = PortPlan(sp.lowered)
    t = g["amplitudes"]
    port.load(0, t, creation=False)
    R = port.apply(0, np.zeros(p.num_amplitudes, complex), sp.diagonal, t)
    assert np.linalg.norm(p.tilde.pack(R, check=False)) <= 1e-12
    # a plausible mistake fails: the same amplitudes with one T2 entry's sign flipped do not
    t_bad = t.copy()
    k = int(np.argmax(np.abs(t)))
    t_bad[k] = -t_bad[k]
    port.load(0, t_bad, creation=False)
    R_bad = port.apply(0, np.zeros(p.num_amplitudes, complex), sp.diagonal, t_bad)
    assert np.linalg.norm(p.tilde.pack(R_bad, check=False)) > 1e-6


def test_s135_action_matches_the_reference_golden_action():
    """The full S1 + S3 + S5 action of the reference's own S135 test (its compact model,
    amplitudes 6e-4 cos(i + 0.19), charge +1, a seeded input; written by
    ``tests/golden/make_config1_reference_vectors.py``) against the compiled S135 sigma program
    run by the NumPy port: ``reference_response_order`` reproduces the reference's coordinate
    order exactly (the S13 part as here, the rank-five shell by key and index combination), and
    the action agrees to 1e-13 relative in that order."""

    from pathlib import Path

    from c3lindblad_amd.configs import s135_validation_model
    from c3lindblad_amd.eom import ComponentEOMProblem, eom_ccsd_problem
    from c3lindblad_amd.lowering import factorized_cc_problem, reference_response_order
    from c3lindblad_amd.static import ComponentStaticProblem
    from oracle import polynomial as poly
    from oracle.port import PortPlan

    g = np.load(Path(__file__).parent / "golden" / "s135_action_reference_vectors.npz")
    p = factorized_cc_problem(s135_validation_model())
    assert poly.coordinate_masks(p.coordinates) == [int(m) for m in g["static_masks"]]
    sp = ComponentStaticProblem.compile(p)
    ep = ComponentEOMProblem.compile(eom_ccsd_problem(p, space="S135"), sp, 1)
    order = reference_response_order(ep.coordinates)
    masks = np.asarray(poly.coordinate_masks(ep.coordinates), dtype=np.int64)
    assert np.array_equal(masks[order], g["response_masks"])
    assert np.array_equal(order[: int(g["p_dimension"])], np.arange(int(g["p_dimension"])))
    x = np.empty(ep.dimension, dtype=np.complex128)
    x[order] = g["x"]
    port = PortPlan(ep.lowered)
    port.load(0, g["amplitudes"])
    port.load(2, x)
    y = port.apply(3, np.zeros(ep.dimension, complex), ep.diagonal, x)
    assert _rel(y[order], g["y"]) < 1e-13

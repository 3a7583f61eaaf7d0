"""Drop-in boundary on the host: the reference's call shapes, result fields and model record.

No GPU is needed: every check here runs before the device is required (keyword handling,
validation) or is pure host code (model records, preconditioner diagnostics, Matsubara grids,
the reduced Krylov solves of ``ArnoldiResolvent``).
"""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"


def _readme_static():
    from c3lindblad_amd.configs import readme_model
    from c3lindblad_amd.lowering import factorized_cc_problem
    from c3lindblad_amd.static import ComponentStaticProblem

    return ComponentStaticProblem.compile(factorized_cc_problem(readme_model()))


def test_solve_component_ccsd_accepts_the_reference_call_shape():
    """``solve_component_ccsd(static, policy=..., rank6_tile_size=..., compute_rank6=False)`` as
    the reference's examples call it (``examples/kanamori/run.py:426``) gets past the signature;
    only the rank-six diagnostic itself is refused."""

    import torch

    from c3lindblad_amd import NewtonKrylovPolicy, solve_component_ccsd

    sp = _readme_static()
    with pytest.raises(NotImplementedError, match="rank-six"):
        solve_component_ccsd(sp, policy=NewtonKrylovPolicy(), compute_rank6=True)
    with pytest.raises(ValueError, match="rank6_tile_size"):
        solve_component_ccsd(sp, rank6_tile_size=0)
    if not torch.cuda.is_available():
        with pytest.raises(RuntimeError, match="GPU"):  # the keywords are accepted; the device is required
            solve_component_ccsd(sp, policy=NewtonKrylovPolicy(), rank6_tile_size=4096, compute_rank6=False)


def test_solution_record_has_the_reference_fields():
    import dataclasses

    from c3lindblad_amd import ComponentCCSolution

    names = {f.name for f in dataclasses.fields(ComponentCCSolution)}
    reference = {"problem", "variables", "amplitudes", "residual", "residual_norm", "packed_residual_norm",
                 "reference_residual", "rank6", "converged", "iterations", "residual_evaluations",
                 "jvp_evaluations", "history", "preconditioner", "wall_seconds", "peak_rss_kib"}
    assert reference <= names


def test_rapidity_preconditioner_diagnostics():
    """``newton_krylov.py:158-192``: threshold 1e-12 max(max|d|, 1), condition max/min."""

    from c3lindblad_amd.solver import rapidity_preconditioner

    sp = _readme_static()
    inverse, diag = rapidity_preconditioner(sp)
    d = -1j * sp.oracle.rapidity_sums()
    mags = np.abs(d)
    assert diag.regularization_threshold == pytest.approx(1e-12 * max(mags.max(), 1.0))
    assert diag.regularized_amplitudes == int(np.count_nonzero(mags < diag.regularization_threshold))
    assert diag.maximum_magnitude == pytest.approx(mags.max())
    assert diag.minimum_magnitude == pytest.approx(mags[mags >= diag.regularization_threshold].min())
    assert diag.magnitude_condition == pytest.approx(diag.maximum_magnitude / diag.minimum_magnitude)
    big = mags >= diag.regularization_threshold
    assert np.allclose(inverse[big] * d[big], 1.0)


def test_reference_model_record_round_trip():
    """A model record in the reference's schema (``models.py:428-452``: ``model_type``,
    ``{"real","imag"}`` matrices, ``interaction_sites``, ``two_body_interaction``) is read by
    ``from_dict`` and written back unchanged by ``to_dict``."""

    from c3lindblad_amd import FermionicLindbladModel

    record = json.loads((GOLDEN / "reference_model_record.json").read_text())
    model = FermionicLindbladModel.from_dict(record)
    assert model.n_sites == 4 and model.interaction_sites == (0, 1) and model.units == "t"
    assert np.allclose(model.E, np.asarray(record["E"]["real"]))
    again = model.to_dict()
    for key in ("model_type", "symmetry", "E", "gamma_loss", "gamma_gain", "U", "impurity_site", "units",
                "input_tolerance", "interaction_sites"):
        assert again[key] == record[key], key
    # canonical terms: zero entries dropped, each pair (p<q, r<s) once, Hermitian partners present
    terms = {tuple(t["indices"]): complex(*t["coefficient"]) for t in again["two_body_interaction"]["terms"]}
    assert (0, 2, 2, 0) not in terms and (0, 2, 0, 2) in terms
    assert terms[(0, 3, 1, 2)] == pytest.approx(-0.2)  # (0,3,2,1) folded in with its sign
    assert FermionicLindbladModel.from_dict(again).to_dict() == again
    with pytest.raises(ValueError, match="model_type"):
        FermionicLindbladModel.from_dict({**record, "model_type": "QuadraticOpenSystem"})


def test_device_amplitudes_must_be_contiguous_complex128():
    """A complex64 or strided tensor would be read as complex128 by the plan: rejected on the host."""

    import torch

    from c3lindblad_amd.eom import _check_device_amplitudes

    with pytest.raises(ValueError):
        _check_device_amplitudes(torch.zeros(6, dtype=torch.complex64), (6,))
    with pytest.raises(ValueError):
        _check_device_amplitudes(torch.zeros(12, dtype=torch.complex128)[::2], (6,))
    with pytest.raises(ValueError):
        _check_device_amplitudes(torch.zeros(6, dtype=torch.complex128), (6,))  # host tensor
    with pytest.raises(ValueError):
        _check_device_amplitudes(np.zeros(6, dtype=complex), (6,))


def test_matsubara_grid_maps_negative_frequencies_to_the_conjugate_transpose():
    from c3lindblad_amd import matsubara_grid

    grid = matsubara_grid(5.0, n_iw=3)
    assert list(grid.indices) == [-3, -2, -1, 0, 1, 2]
    assert np.allclose(grid.upper_shifts, 1j * np.pi / 5.0 * np.array([1, 3, 5]))
    values = np.arange(3 * 4, dtype=float).reshape(3, 2, 2) + 1j
    out = grid.from_upper(values)
    assert np.allclose(out[3], values[0]) and np.allclose(out[2], values[0].conj().T)
    assert np.allclose(out[0], values[2].conj().T)
    same = matsubara_grid(5.0, omega_n=grid.omega_n)
    assert np.array_equal(same.indices, grid.indices)
    with pytest.raises(ValueError):
        matsubara_grid(5.0, omega_n=np.array([2 * np.pi / 5.0]))  # bosonic point


def test_arnoldi_resolvent_reduced_solves_on_the_host():
    """The reduced (Schur) solves of ``ArnoldiResolvent`` on an exact Arnoldi factorisation of a
    dense non-normal matrix built on the CPU: coefficients equal ``(z - H_m)^{-1} beta e_1`` and the
    residual norms equal the true residuals of ``V_m y``."""

    import torch

    from c3lindblad_amd.response import ArnoldiResolvent
    from oracle import solvers

    rng = np.random.default_rng(5)
    n = 40
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) + 4.0 * np.diag(rng.normal(size=n))
    b = rng.normal(size=n) + 1j * rng.normal(size=n)
    V, H, beta = solvers.arnoldi_resolvent(lambda v: A @ v, b, np.array([0.3j]), maximum_dimension=25,
                                           tolerance=1e-300, minimum_dimension=25)[:3]
    m = H.shape[1]
    proj = ArnoldiResolvent(basis=torch.from_numpy(np.ascontiguousarray(V.T)), hessenberg=H, source_norm=beta,
                            breakdown=False, shifts=np.array([0.3j]), green=np.zeros(1, complex),
                            shift_relative_residuals=np.zeros(1), converged=True, construction_wall_seconds=0.0,
                            action_calls=m)
    z = np.array([1.0 + 0.5j, -2.0 + 0.1j, 3.0j])
    y = proj.coefficients(z)
    e1 = np.zeros(m, dtype=complex)
    e1[0] = beta
    for zk, yk in zip(z, y):
        assert np.allclose(yk, np.linalg.solve(zk * np.eye(m) - H[:m, :m], e1), rtol=1e-10, atol=1e-12)
        x = V[:, :m] @ yk
        true = np.linalg.norm(b - (zk * np.eye(n) - A) @ x) / beta
        assert true == pytest.approx(proj.relative_residuals(np.array([zk]))[0], rel=1e-6, abs=1e-12)
    t = np.array([0.0, 0.4])
    c = proj.time_coefficients(t)
    import scipy.linalg

    assert np.allclose(c[1], scipy.linalg.expm(-0.4j * H[:m, :m]) @ e1, rtol=1e-10, atol=1e-12)
    assert np.allclose(c[0], e1)


def test_checkpoint_resource_guard_and_identity():
    from c3lindblad_amd.checkpoint import require_resources

    require_resources(required_memory_bytes=10, required_disk_bytes=10, available_memory_bytes=10, free_disk_bytes=10)
    with pytest.raises(MemoryError):
        require_resources(required_memory_bytes=11, required_disk_bytes=0, available_memory_bytes=10, free_disk_bytes=0)
    with pytest.raises(OSError):
        require_resources(required_memory_bytes=0, required_disk_bytes=11, available_memory_bytes=0, free_disk_bytes=10)


def test_batch_skips_instances_without_a_full_hf_reference():
    """The reference raises for a model whose default full-HF reference does not converge
    (``cc/factorized.py:304-318``); the batch set-up leaves such an instance out and names it."""

    from c3lindblad_amd.batch import _compile_job
    from c3lindblad_amd.configs import multiorbital_model, siam_model

    bad = multiorbital_model(3, 1, seed=0)
    inst, why = _compile_job((7, bad, (1,)))
    assert inst is None and why[0] == 7 and "full-HF" in why[1]
    inst, why = _compile_job((3, siam_model(3, 2.0, seed=1), (1, -1)))
    assert why is None and inst.index == 3 and set(inst.eom) == {1, -1}


def test_s135_space_keys_and_layout():
    """S135 adds the rank-five keys of the charge (``cc/eom.py:1-40``) after the S13 strings, which
    keep their S13 order, so S13 vectors embed as a prefix; unknown spaces are rejected."""

    from math import comb

    from c3lindblad_amd.eom import eom_ccsd_problem
    from c3lindblad_amd.lowering import factorized_cc_problem, response_coordinates, response_keys

    assert response_keys(1, "S135")[2] == [(3, 2, 0, 0), (2, 1, 1, 1), (1, 0, 2, 2)]
    assert response_keys(-1, "S135")[2] == [(2, 3, 0, 0), (1, 2, 1, 1), (0, 1, 2, 2)]
    for n in (2, 3, 5):
        for q in (1, -1):
            s13 = response_coordinates(n, q)
            s135 = response_coordinates(n, q, "S135")
            shell = comb(n, 3) * comb(n, 2) + comb(n, 2) * n**3 + n * comb(n, 2) ** 2
            assert s135.dimension == s13.dimension + shell
            width = s13.modes.shape[1]
            assert np.array_equal(s135.modes[: s13.dimension, :width], s13.modes)
            assert np.all(s135.modes[: s13.dimension, width:] == -1)
    with pytest.raises(ValueError, match="space"):
        response_keys(1, "S7")
    cc = _readme_static().oracle
    with pytest.raises(ValueError, match="space"):
        eom_ccsd_problem(cc, space="S1357")
    problem = eom_ccsd_problem(cc, space="S135")
    assert problem.max_rank == 5 and problem.dimension == sum(
        response_coordinates(3, q, "S135").dimension for q in (1, -1))


def test_s135_reference_knobs_are_validated_before_the_device():
    from types import SimpleNamespace

    from c3lindblad_amd.s135 import CompiledFullS135Action, FullS135ResponseOperator

    with pytest.raises(ValueError, match="materialization_batch_size"):
        CompiledFullS135Action(SimpleNamespace(), materialization_batch_size=0)
    with pytest.raises(ValueError, match="s5_tile_size"):
        FullS135ResponseOperator(SimpleNamespace(), s5_tile_size=0)


def test_residual_degree_is_four():
    """The residual is a quartic polynomial of the amplitudes (two-body products), which the
    two-node Gauss-Legendre residual path relies on."""

    assert _readme_static().residual_degree == 4

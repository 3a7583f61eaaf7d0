"""Writes golden vectors from the reference's own NumPy path.

Run in the build container only (the reference is not on the GPU box and never travels): it
imports the reference from ``/root/reference/src``.  The fixtures pin what the brute-force and
closed-form checks cannot: converged amplitudes in the reference's own gauge (full-HF
reference, normal-mode canonicalisation) and coordinate order, keyed by their particle-hole
mode masks, and the reference's full S1 + S3 + S5 action in its own coordinate order.

* ``config1``: the README quick-start model (3 orbitals, spin 1/2, U = 1.2 on orbital 0; CCSD
  steady state, then the charge +1 EOM-CCSD retarded Green function on 81 frequencies in
  [-4, 4], eta = 0) solved exactly as the reference README does, plus the charge -1 Green
  function and the occupations (covariance) -> ``config1_reference_vectors.npz``;
* ``families``: two n = 6 models of the benchmark families, built by this repository's
  generators and handed to the reference through its own model record (to_dict / from_dict):
  a SIAM with five pseudomodes at U = 4 and two Kanamori orbitals with V_pqrs and two
  pseudomodes each -> ``siam6_reference_vectors.npz``, ``multiorbital6_reference_vectors.npz``;
* ``s135``: the full S1 + S3 + S5 response action of the reference's own S135 test
  (``tests/cc/test_s135_full.py``: its compact model, amplitudes 6e-4 cos(i + 0.19), charge +1,
  a seeded complex input) -> ``s135_action_reference_vectors.npz``.

    python tests/golden/make_config1_reference_vectors.py [config1] [families] [s135]

(no argument: all three).
"""

from __future__ import annotations

import sys
import types
from pathlib import Path

import numpy as np

REFERENCE_SRC = Path("/root/reference/src")
HERE = Path(__file__).resolve().parent
CASES = ("config1", "families", "s135")


def _solve_and_save(ref, model, omega, out: Path, removal_charge: bool = False) -> None:
    """CCSD root, occupations (covariance) and charge +1 G^R of one model by the reference (and
    the charge -1 G^R when asked); saves the fixture."""

    oracle = ref.factorized_cc_problem(model, "CCSD")
    problem = ref.ComponentStaticProblem.compile(oracle)
    solution = ref.solve_component_ccsd(problem)
    eom = ref.eom_ccsd_problem(problem.oracle)
    green = {}
    for charge in ((1, -1) if removal_charge else (1,)):
        context = ref.component_eom_context(ref.ComponentEOMProblem.compile(eom, problem, charge=charge),
                                            solution.amplitudes)
        green[charge] = ref.component_retarded_green(ref.ComponentSigmaOperator(context),
                                                     ref.component_eom_sources(context), omega, eta=0.0)
    covariance = ref.factorized_ccsd_covariance(oracle.reference, oracle.cluster_basis, solution.amplitudes)
    extra = {}
    if removal_charge:  # both charges, each channel on its own
        extra = {"retarded_minus": np.asarray(green[-1].retarded, dtype=np.complex128),
                 "addition_plus": np.asarray(green[1].addition, dtype=np.complex128),
                 "removal_plus": np.asarray(green[1].removal, dtype=np.complex128),
                 "addition_minus": np.asarray(green[-1].addition, dtype=np.complex128),
                 "removal_minus": np.asarray(green[-1].removal, dtype=np.complex128)}
    response = types.SimpleNamespace(retarded=green[1].retarded)
    np.savez(
        out,
        **extra,
        covariance=np.asarray(covariance, dtype=np.complex128),
        omega=omega,
        retarded=np.asarray(response.retarded, dtype=np.complex128),
        amplitudes=np.asarray(solution.amplitudes, dtype=np.complex128),
        masks=np.asarray(oracle.cluster_basis.masks, dtype=np.int64),
        rapidities=np.asarray(oracle.rapidities, dtype=np.complex128),
        packed_residual_norm=np.float64(solution.packed_residual_norm),
    )
    print(f"wrote {out}: {len(oracle.cluster_basis.masks)} amplitudes, packed residual "
          f"{solution.packed_residual_norm:.2e}, G^R shape {np.shape(response.retarded)}")


def main(cases) -> None:
    sys.path.insert(0, str(REFERENCE_SRC))
    sys.path.insert(1, str(HERE.parents[1]))  # this repository, for its model generators

    from c3l import FermionicLindbladModel
    from c3l.cc.component_response import ComponentSigmaOperator, component_retarded_green
    from c3l.cc.component_solver import solve_component_ccsd
    from c3l.cc.component_tensor import (ComponentEOMProblem, ComponentStaticProblem, component_eom_context,
                                         component_eom_sources)
    from c3l.cc.eom import eom_ccsd_problem
    from c3l.cc.factorized import factorized_cc_problem
    from c3l.cc.observables import factorized_ccsd_covariance

    ref = types.SimpleNamespace(
        factorized_ccsd_covariance=factorized_ccsd_covariance,
        factorized_cc_problem=factorized_cc_problem, ComponentStaticProblem=ComponentStaticProblem,
        solve_component_ccsd=solve_component_ccsd, eom_ccsd_problem=eom_ccsd_problem,
        component_eom_context=component_eom_context, ComponentEOMProblem=ComponentEOMProblem,
        component_retarded_green=component_retarded_green, ComponentSigmaOperator=ComponentSigmaOperator,
        component_eom_sources=component_eom_sources)

    if "config1" in cases:
        model = FermionicLindbladModel(
            E=np.array([[-0.6, 0.28, 0.28], [0.28, -1.2, 0.0], [0.28, 0.0, 1.2]]),
            gamma_loss=np.diag([0.0, 0.06, 0.48]),
            gamma_gain=np.diag([0.0, 0.48, 0.06]),
            U=1.2,
        )
        _solve_and_save(ref, model, np.linspace(-4.0, 4.0, 81), HERE / "config1_reference_vectors.npz",
                        removal_charge=True)

    if "families" in cases:
        from c3lindblad_amd.configs import multiorbital_model, siam_model

        for name, ours in (("siam6", siam_model(5, 4.0, seed=3)), ("multiorbital6", multiorbital_model(2, 2, seed=3))):
            _solve_and_save(ref, FermionicLindbladModel.from_dict(ours.to_dict()), np.linspace(-6.0, 6.0, 61),
                            HERE / f"{name}_reference_vectors.npz")

    if "s135" in cases:
        from c3l.cc.s135_full import CompiledFullS135Action, FullS135ResponseOperator
        from c3lindblad_amd.configs import s135_validation_model

        model = FermionicLindbladModel.from_dict(s135_validation_model().to_dict())
        oracle = factorized_cc_problem(model, "CCSD")
        static = ComponentStaticProblem.compile(oracle)
        coordinate = np.arange(oracle.num_variables, dtype=np.float64)
        amplitudes = oracle.tilde.unpack(6.0e-4 * np.cos(coordinate + 0.19))
        eom = eom_ccsd_problem(oracle, response_spin=0, build_next_rank_masks=False)
        context = component_eom_context(ComponentEOMProblem.compile(eom, static, 1), amplitudes)
        operator = FullS135ResponseOperator(CompiledFullS135Action(context, materialization_batch_size=4),
                                            s5_tile_size=7)
        rng = np.random.default_rng(20260722)
        x = rng.normal(size=operator.dimension) + 1j * rng.normal(size=operator.dimension)
        y = operator.matvec(x)
        coords = operator.coordinates
        masks = np.zeros(coords.dimension, dtype=np.int64)
        for block in coords.blocks:
            groups = [g for g, count in enumerate(block.key) for _ in range(count)]
            for position, indices in zip(block.vector_positions, block.indices):
                masks[int(position)] = sum(1 << (g * coords.n + int(i)) for g, i in zip(groups, indices))
        out = HERE / "s135_action_reference_vectors.npz"
        np.savez(out, amplitudes=np.asarray(amplitudes, dtype=np.complex128),
                 static_masks=np.asarray(oracle.cluster_basis.masks, dtype=np.int64), response_masks=masks, x=x,
                 y=np.asarray(y, dtype=np.complex128), p_dimension=np.int64(operator.p_dimension))
        print(f"wrote {out}: response dimension {operator.dimension} (S13 part {operator.p_dimension})")


if __name__ == "__main__":
    chosen = tuple(sys.argv[1:]) or CASES
    unknown = set(chosen) - set(CASES)
    if unknown:
        raise SystemExit(f"unknown cases {sorted(unknown)}; choose from {CASES}")
    main(chosen)

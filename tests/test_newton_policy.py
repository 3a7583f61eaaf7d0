"""Host logic of the Newton--Krylov policy: forcing terms (no GPU)."""

from __future__ import annotations

import pytest

from c3lindblad_amd.solver import NewtonKrylovPolicy, forcing_tolerance


def test_legacy_forcing_is_a_tenth_of_the_residual_within_the_bounds():
    p = NewtonKrylovPolicy()
    assert forcing_tolerance(1e-4, p) == pytest.approx(1e-5)
    assert forcing_tolerance(10.0, p) == p.maximum_gmres_relative_tolerance
    assert forcing_tolerance(1e-12, p) == p.minimum_gmres_relative_tolerance


def test_eisenstat_walker_choice_two_with_safeguard():
    p = NewtonKrylovPolicy(forcing_strategy="eisenstat_walker", maximum_gmres_relative_tolerance=0.9,
                           initial_gmres_relative_tolerance=0.5, eisenstat_walker_exponent=2.0)
    assert forcing_tolerance(1.0, p) == 0.5  # eta_0
    # gamma (|F_k| / |F_k-1|)^alpha
    assert forcing_tolerance(1e-2, p, previous_norm=1e-1, previous_tolerance=1e-3) == pytest.approx(0.9 * 1e-2)
    # safeguard: gamma eta_{k-1}^alpha when it is larger
    assert forcing_tolerance(1e-3, p, previous_norm=1e-1, previous_tolerance=0.5) == pytest.approx(0.9 * 0.25)
    assert forcing_tolerance(1e-12, p, previous_norm=1.0, previous_tolerance=1e-9) == p.minimum_gmres_relative_tolerance


def test_mixed_precision_policy_constants():
    legacy = NewtonKrylovPolicy.mixed_precision()
    assert legacy.jacobian_precision == "fp32" and legacy.minimum_gmres_relative_tolerance == 1e-6
    assert legacy.maximum_gmres_relative_tolerance == NewtonKrylovPolicy().maximum_gmres_relative_tolerance
    ew = NewtonKrylovPolicy.mixed_precision(forcing_strategy="eisenstat_walker")
    assert (ew.maximum_gmres_relative_tolerance, ew.initial_gmres_relative_tolerance, ew.eisenstat_walker_gamma,
            ew.eisenstat_walker_exponent) == (0.9, 0.5, 0.9, 2.0)
    tuned = NewtonKrylovPolicy.mixed_precision(forcing_strategy="eisenstat_walker", eisenstat_walker_exponent=1.5)
    assert tuned.eisenstat_walker_exponent == 1.5  # explicit overrides win
    with pytest.raises(ValueError):
        NewtonKrylovPolicy.mixed_precision(minimum_gmres_relative_tolerance=1.0)  # floor above the cap

"""Host lowering: coordinate order, dimensions, tilde structure, validation (no GPU)."""

from __future__ import annotations

from itertools import combinations
from math import comb

import numpy as np
import pytest

from tests.conftest import random_model


def _mask(modes):
    return sum(1 << int(m) for m in modes)


@pytest.mark.parametrize("n", [1, 2, 3])
def test_static_coordinates_follow_reference_mask_order(n):
    """Strings sorted by (PH rank, integer mask) as ``ci/basis.py:149-178`` builds them."""

    from c3lindblad_amd.lowering import static_coordinates

    cs = static_coordinates(n)
    masks = [_mask(r[r >= 0]) for r in cs.modes]
    groups = [range(g * n, (g + 1) * n) for g in range(4)]
    expected = []
    for rank in (1, 2):
        for a in range(rank + 1):
            b = rank - a
            for up in combinations(groups[0], a):
                for um in combinations(groups[1], a):
                    for dp in combinations(groups[2], b):
                        for dm in combinations(groups[3], b):
                            expected.append((rank, _mask(up + um + dp + dm)))
    expected = [m for _, m in sorted(expected)]
    assert masks == expected
    assert cs.dimension == 2 * n**2 + 2 * comb(n, 2) ** 2 + n**4


@pytest.mark.parametrize("n,charge", [(2, 1), (3, -1)])
def test_response_coordinates_follow_reference_order(n, charge):
    from c3lindblad_amd.lowering import response_coordinates

    cs = response_coordinates(n, charge)
    masks = [_mask(r[r >= 0]) for r in cs.modes]
    ranks = [(int(d) - 1) // 2 for d in cs.degrees]
    assert list(zip(ranks, masks)) == sorted(zip(ranks, masks))
    assert cs.dimension == n + n * comb(n, 2) + n**3
    charges = []
    for r in cs.modes:
        q = 0
        for m in r[r >= 0]:
            g = m // n
            assert g in (0, 1, 2, 3)
            q += (1 if g == 0 else -1 if g == 1 else 0)
        charges.append(q)
    assert set(charges) == {charge}


def test_block_positions_cover_the_vector_once():
    from c3lindblad_amd.lowering import static_coordinates

    cs = static_coordinates(3)
    pos = np.concatenate([b.positions for b in cs.blocks])
    assert np.array_equal(np.sort(pos), np.arange(cs.dimension))
    for b in cs.blocks:
        modes = cs.modes[b.positions, : b.degree]
        slots = [g for g, c in enumerate(b.key) for _ in range(c)]
        assert np.array_equal(modes, b.indices + np.array(slots) * 3)


def test_tilde_map_is_an_involution_with_consistent_phases():
    from c3lindblad_amd.lowering import factorized_cc_problem

    p = factorized_cc_problem(random_model(3, 1.1, 4))
    tl = p.tilde
    assert np.array_equal(tl.partner[tl.partner], np.arange(tl.dimension))
    assert np.max(np.abs(tl.phase[tl.partner] * tl.phase.conj() - 1.0)) < 1e-10
    assert tl.real_size == tl.dimension
    x = np.random.default_rng(0).normal(size=tl.real_size)
    t = tl.unpack(x)
    assert tl.defect(t) < 1e-14
    assert np.allclose(tl.pack(t), x, atol=1e-14)


def test_siam_products_are_interaction_pairs_plus_impurity_counterterms():
    from c3lindblad_amd.configs import siam_model
    from c3lindblad_amd.lowering import factorized_cc_problem

    p = factorized_cc_problem(siam_model(4, 2.0, seed=1))
    labels = [r.label for r in p.products]
    assert sum(lbl.startswith("two_body") for lbl in labels) == 2
    assert all("_0_" in lbl for lbl in labels if lbl.startswith("quadratic_counter"))
    assert len(p.products) == 6


@pytest.mark.parametrize(
    "kwargs,error",
    [
        (dict(E=[[0.0, 1.0], [0.0, 0.0]]), ValueError),
        (dict(E=[[0.0]], gamma_loss=[[-1.0]]), ValueError),
        (dict(E=[[np.nan]]), ValueError),
        (dict(E=[[0.0]], U=1j), ValueError),
        (dict(E=[[0.0]], impurity_site=3), ValueError),
    ],
)
def test_model_validation_matches_reference_errors(kwargs, error):
    from c3lindblad_amd.model import FermionicLindbladModel

    n = np.asarray(kwargs["E"]).shape[0]
    base = dict(gamma_loss=np.zeros((n, n)), gamma_gain=np.zeros((n, n)), U=0.0)
    base.update(kwargs)
    with pytest.raises(error):
        FermionicLindbladModel(**base)

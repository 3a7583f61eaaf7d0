"""Pins of the NumPy restatement of the warm-start MLP (no GPU).

* ``backward`` (exact float64 mode) against central finite differences of the
  loss, for every weight and bias entry of a small network;
* ``bf16`` rounding: exactly representable values stay, ties go to even,
  the bit patterns equal the top 16 bits of the rounded float32.
"""

from __future__ import annotations

import numpy as np


def _net(rng, sizes):
    W = [rng.normal(scale=0.5, size=(k, n)) for k, n in zip(sizes[:-1], sizes[1:])]
    b = [rng.normal(scale=0.1, size=n) for n in sizes[1:]]
    return W, b


def test_mlp_backward_equals_finite_differences():
    from oracle import mlp

    rng = np.random.default_rng(5)
    sizes = (7, 6, 5, 6, 3)
    W, b = _net(rng, sizes)
    x = rng.normal(size=(4, sizes[0]))
    t = rng.normal(size=(4, sizes[-1]))
    _, gW, gb = mlp.loss_and_backward(W, b, x, t, exact=True)
    h = 1e-6
    for l in range(len(W)):
        for idx in np.ndindex(W[l].shape):
            Wp = [w.copy() for w in W]
            Wm = [w.copy() for w in W]
            Wp[l][idx] += h
            Wm[l][idx] -= h
            fd = (mlp.loss_and_backward(Wp, b, x, t, exact=True)[0] - mlp.loss_and_backward(Wm, b, x, t, exact=True)[0]) / (2 * h)
            assert abs(fd - gW[l][idx]) < 1e-7 * max(1.0, abs(fd))
        for i in range(b[l].size):
            bp = [v.copy() for v in b]
            bm = [v.copy() for v in b]
            bp[l][i] += h
            bm[l][i] -= h
            fd = (mlp.loss_and_backward(W, bp, x, t, exact=True)[0] - mlp.loss_and_backward(W, bm, x, t, exact=True)[0]) / (2 * h)
            assert abs(fd - gb[l][i]) < 1e-7 * max(1.0, abs(fd))


def test_bf16_rounding():
    from oracle import mlp

    exact = np.array([1.0, -2.5, 0.15625, 2.0**100], dtype=np.float32)
    assert np.array_equal(mlp.bf16(exact), exact)
    # 1 + 2^-8 lies halfway between 1 and 1 + 2^-7: ties to even -> 1
    assert mlp.bf16(np.array([1.0 + 2.0**-8], dtype=np.float32))[0] == 1.0
    # 1 + 3 * 2^-8 is halfway between 1 + 2^-7 and 1 + 2^-6: ties to even -> 1 + 2^-6
    assert mlp.bf16(np.array([1.0 + 3 * 2.0**-8], dtype=np.float32))[0] == np.float32(1.0 + 2.0**-6)
    v = mlp.bf16(np.random.default_rng(1).normal(size=100).astype(np.float32))
    assert np.array_equal((v.view(np.uint32) & 0xFFFF), np.zeros(100, dtype=np.uint32))
    assert np.array_equal(mlp.bits(v).astype(np.uint32) << 16, v.view(np.uint32))

"""NumPy restatement of the warm-start MLP -- TEST INFRASTRUCTURE ONLY.

Same network and the same rounding points as the HIP kernels
(``c3lindblad_amd/csrc/mlp.hip``): layer inputs and weights in bfloat16
(round to nearest even), products accumulated in FP32, biases and
pre-activations FP32, GELU in its tanh form, MSE loss
``sum (y - t)^2 / (B * outputs)``.  With ``exact=True`` every rounding is
skipped and everything runs in float64, which is what the finite-difference
pin of the backward formulas uses.

The learned warm start has no counterpart in the reference; its pins are the
finite-difference check of ``backward`` (``tests/test_oracle_pins.py``) and
the GPU parity tests of the kernels against this file.
"""

from __future__ import annotations

import numpy as np

C0 = 0.7978845608028654  # sqrt(2 / pi)
C1 = 0.044715


def bf16(a: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float32."""

    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, (u | 0x400000) & 0xFFFF0000, r)
    return r.astype(np.uint32).view(np.float32).reshape(np.shape(a))


def bits(a: np.ndarray) -> np.ndarray:
    """bfloat16 bit patterns (uint16) of already-rounded float32 values."""

    return (np.ascontiguousarray(a, dtype=np.float32).view(np.uint32) >> 16).astype(np.uint16)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(C0 * (z + C1 * z**3)))


def gelu_grad(z):
    th = np.tanh(C0 * (z + C1 * z**3))
    return 0.5 * (1.0 + th) + 0.5 * z * (1.0 - th * th) * C0 * (1.0 + 3.0 * C1 * z * z)


def forward(weights, biases, x, exact: bool = False):
    """Return ``(y, cache)``; ``weights[l]`` is ``K_l x N_l`` (y = x W + b)."""

    if exact:
        h = np.asarray(x, dtype=np.float64)
    else:
        h = bf16(np.asarray(x, dtype=np.float32))
    cache = {"h": [h], "z": []}
    L = len(weights)
    for l, (W, b) in enumerate(zip(weights, biases)):
        if exact:
            z = h @ W + b
        else:
            z = (h.astype(np.float64) @ bf16(W).astype(np.float64)).astype(np.float32) + np.float32(b)
            z = z.astype(np.float32)
        cache["z"].append(z)
        if l < L - 1:
            a = gelu(z)
            h = a if exact else bf16(a.astype(np.float32))
            cache["h"].append(h)
        else:
            h = z
    return h, cache


def loss_and_backward(weights, biases, x, target, exact: bool = False):
    """MSE loss and its gradients with respect to every weight and bias."""

    y, cache = forward(weights, biases, x, exact)
    B, n_out = y.shape
    d = y - target
    loss = float(np.sum(np.asarray(d, dtype=np.float64) ** 2) / (B * n_out))
    dy = (2.0 / (B * n_out)) * d
    if not exact:
        dy = dy.astype(np.float32)
    L = len(weights)
    gw, gb = [None] * L, [None] * L
    g = dy
    for l in range(L - 1, -1, -1):
        gq = g if exact else bf16(g)
        h = cache["h"][l]
        if exact:
            gw[l] = h.T @ g
        else:
            gw[l] = (h.astype(np.float64).T @ gq.astype(np.float64)).astype(np.float32)
        gb[l] = np.sum(g, axis=0) if exact else np.sum(g.astype(np.float64), axis=0).astype(np.float32)
        if l > 0:
            if exact:
                dh = g @ weights[l].T
                g = dh * gelu_grad(cache["z"][l - 1])
            else:
                dh = (gq.astype(np.float64) @ bf16(weights[l]).astype(np.float64).T).astype(np.float32)
                g = (dh * gelu_grad(cache["z"][l - 1])).astype(np.float32)
    return loss, gw, gb


def adam(w, m, v, g, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    w = w - lr * (m / (1 - b1**step)) / (np.sqrt(v / (1 - b2**step)) + eps)
    return w, m, v

"""Learned warm start for the CCSD Newton solve (no counterpart in the reference).

An MLP maps model parameters to initial amplitudes in a compressed basis:

* features: the flattened ``(E, γ_loss, γ_gain)`` matrices padded to 32 sites
  (real and imaginary part of ``E``) and ``U``, with the non-impurity sites sorted
  by their level so that equivalent models give equal features, standardised with
  the training set's mean and spread;
* targets: the converged amplitudes in a canonical mode labelling (within each of
  the four particle-hole mode groups the mode with the largest impurity weight
  first, then increasing ``Re λ``; every string re-sorted with its fermionic sign), as
  ``[Re t, Im t]``, expressed in the leading right singular vectors of the centred
  training matrix (rank 64 by default) and scaled to unit spread;
* network: 3 hidden layers of width 512 with GELU, bfloat16 weights and layer
  inputs, FP32 accumulation on the matrix cores (``c3l_bf16_gemm``), FP32
  master weights trained with Adam (``c3l_mlp_adam``), MSE loss;
* blend: the prediction is scaled by the factor ``α`` that minimises the amplitude
  error on a held-out part of the training solutions (``α -> 0`` recovers the zero
  guess when the network generalises poorly).

``initial_variables(problem)`` maps ``mean + basis @ coefficients`` back to the
problem's own mode labelling and packs it into real tilde variables;
``solve_component_ccsd(..., initial_variables=...)`` iterates from it to the same
residual tolerance and restarts from the zero guess if Newton does not converge
from it.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib

PAD = 32
HIDDEN = (512, 512, 512)


def model_features(model, pad: int = PAD) -> np.ndarray:
    """Flattened ``(E, γ_loss, γ_gain, U)`` padded to ``pad`` sites (float32).

    The impurity site comes first and the other sites follow in increasing level
    ``Re E[k, k]`` (a relabelling of sites leaves the physics unchanged).
    """

    n = model.n_sites
    if n > pad:
        raise ValueError(f"models above {pad} sites are outside the feature layout")
    E = np.asarray(model.E)
    imp = int(getattr(model, "impurity_site", 0))
    rest = [k for k in range(n) if k != imp]
    rest.sort(key=lambda k: (float(np.real(E[k, k])), k))
    order = np.array([imp] + rest)
    E = E[np.ix_(order, order)]
    parts = []
    for m in (E.real, E.imag if np.iscomplexobj(E) else None, np.asarray(model.gamma_loss)[np.ix_(order, order)].real,
              np.asarray(model.gamma_gain)[np.ix_(order, order)].real):
        block = np.zeros((pad, pad))
        if m is not None:
            block[:n, :n] = m
        parts.append(block.ravel())
    U = getattr(model, "U", 0.0)
    parts.append(np.array([float(np.real(U)) if np.isscalar(U) else 0.0]))
    return np.concatenate(parts).astype(np.float32)


def canonical_order(problem) -> tuple[np.ndarray, np.ndarray]:
    """``(perm, sign)`` with ``canonical[perm[i]] = sign[i] * t[i]`` for the amplitudes ``t``.

    Within each particle-hole mode group the mode with the largest weight on the
    impurity site comes first and the others follow in increasing ``Re λ``; each
    creator string is re-sorted in the new labels and picks up the sign of that sort.
    """

    from .lowering import _order_key

    n = problem.model.n_sites
    imp = int(getattr(problem.model, "impurity_site", 0))
    coords = problem.coordinates
    rap = np.asarray(problem.rapidities)
    label = np.empty(4 * n, dtype=np.int64)
    for spin, ref in enumerate(problem.reference.mode_references):
        weight = np.abs(ref.V[imp, :]) ** 2 + np.abs(ref.V[n + imp, :]) ** 2
        for half in range(2):
            local = np.arange(half * n, (half + 1) * n)
            first = int(local[np.argmax(weight[local])])
            rest = sorted((int(m) for m in local if m != first), key=lambda m: (rap[2 * n * spin + m].real, m))
            g = 2 * spin + half
            label[2 * n * spin + np.array([first] + rest)] = np.arange(g * n, (g + 1) * n)
    dim = coords.dimension
    perm = np.empty(dim, dtype=np.int64)
    sign = np.ones(dim)
    for deg in np.unique(coords.degrees):
        sel = np.flatnonzero(coords.degrees == deg)
        rows = coords.modes[sel, :deg]
        mapped = label[rows]
        inversions = np.zeros(sel.size, dtype=np.int64)
        for a in range(deg):
            for b in range(a + 1, deg):
                inversions += mapped[:, a] > mapped[:, b]
        keys = _order_key(rows)
        target = _order_key(np.sort(mapped, axis=1))
        loc = np.searchsorted(keys, target)
        if np.any(loc >= keys.size) or np.any(keys[np.minimum(loc, keys.size - 1)] != target):
            raise ValueError("the relabelled strings are not in the coordinate set")
        perm[sel] = sel[loc]
        sign[sel] = np.where(inversions % 2 == 1, -1.0, 1.0)
    return perm, sign


def _canonical_real(problem, amplitudes) -> np.ndarray:
    perm, sign = canonical_order(problem)
    c = np.empty(perm.size, dtype=np.complex128)
    c[perm] = sign * np.asarray(amplitudes, dtype=np.complex128)
    return np.concatenate((c.real, c.imag))


def _from_canonical_real(problem, r) -> np.ndarray:
    perm, sign = canonical_order(problem)
    d = perm.size
    c = r[:d] + 1j * r[d:]
    return sign * c[perm]


def _bf16_bits(a) -> np.ndarray:
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    r = np.where(nan, ((u >> 16) | 0x40), r)
    return r.astype(np.uint16).reshape(np.shape(a))


class AmplitudeMLP:
    """The network on the GPU: FP32 master weights, bf16 copies, Adam moments."""

    def __init__(self, sizes, seed: int = 0):
        torch = _lib.require_device()
        self.torch = torch
        self.sizes = tuple(int(s) for s in sizes)
        rng = np.random.default_rng(seed)
        dev = _lib.DEVICE
        self.W, self.b, self.W16 = [], [], []
        for k, n in zip(self.sizes[:-1], self.sizes[1:]):
            w = rng.normal(scale=np.sqrt(2.0 / k), size=(k, n)).astype(np.float32)
            self.W.append(torch.from_numpy(w).to(dev))
            self.b.append(torch.zeros(n, dtype=torch.float32, device=dev))
            self.W16.append(torch.from_numpy(_bf16_bits(w).view(np.int16)).to(dev))
        self.mW = [torch.zeros_like(w) for w in self.W]
        self.vW = [torch.zeros_like(w) for w in self.W]
        self.mb = [torch.zeros_like(b) for b in self.b]
        self.vb = [torch.zeros_like(b) for b in self.b]
        self.step_count = 0

    # -- helpers ---------------------------------------------------------------
    def _stream(self):
        return _lib.stream_ptr(self.torch)

    def set_weights(self, weights, biases) -> None:
        torch = self.torch
        for l, (w, b) in enumerate(zip(weights, biases)):
            w = np.ascontiguousarray(w, dtype=np.float32)
            self.W[l].copy_(torch.from_numpy(w))
            self.b[l].copy_(torch.from_numpy(np.ascontiguousarray(b, dtype=np.float32)))
            self.W16[l].copy_(torch.from_numpy(_bf16_bits(w).view(np.int16)))

    def forward(self, x16, batch: int):
        """``x16``: bf16 bit patterns [batch, sizes[0]] (torch int16).  Returns y and the cache."""

        torch = self.torch
        h = x16
        zs, hs = [], [x16]
        L = len(self.W)
        for l in range(L):
            K, N = self.sizes[l], self.sizes[l + 1]
            z = torch.empty(batch, N, dtype=torch.float32, device=_lib.DEVICE)
            _lib.call("c3l_bf16_gemm", _lib.dptr(h), K, 1, _lib.dptr(self.W16[l]), N, 1, _lib.dptr(z), N, batch, N, K,
                      1.0, 0.0, self._stream())
            last = l == L - 1
            hn = None if last else torch.empty(batch, N, dtype=torch.int16, device=_lib.DEVICE)
            _lib.call("c3l_mlp_bias_act", _lib.dptr(z), _lib.dptr(self.b[l]), None if last else _lib.dptr(hn), batch, N,
                      0 if last else 1, self._stream())
            zs.append(z)
            if not last:
                hs.append(hn)
                h = hn
        return zs[-1], (zs, hs)

    def backward(self, cache, dy, dy16, batch: int):
        """Gradients of every weight and bias from ``dy`` (FP32) / ``dy16`` (bf16 bits)."""

        torch = self.torch
        zs, hs = cache
        L = len(self.W)
        gW = [None] * L
        gb = [None] * L
        g, g16 = dy, dy16
        for l in range(L - 1, -1, -1):
            K, N = self.sizes[l], self.sizes[l + 1]
            gW[l] = torch.empty(K, N, dtype=torch.float32, device=_lib.DEVICE)
            # dW = h^T g: A(m=k, k=b) = h[b, k]
            _lib.call("c3l_bf16_gemm", _lib.dptr(hs[l]), 1, K, _lib.dptr(g16), N, 1, _lib.dptr(gW[l]), N, K, N, batch,
                      1.0, 0.0, self._stream())
            gb[l] = torch.empty(N, dtype=torch.float32, device=_lib.DEVICE)
            _lib.call("c3l_mlp_column_sum", _lib.dptr(g), _lib.dptr(gb[l]), batch, N, self._stream())
            if l > 0:
                dh = torch.empty(batch, K, dtype=torch.float32, device=_lib.DEVICE)
                # dh = g W^T: B(k=n, n=k') = W[k', n]
                _lib.call("c3l_bf16_gemm", _lib.dptr(g16), N, 1, _lib.dptr(self.W16[l]), 1, N, _lib.dptr(dh), K, batch, K,
                          N, 1.0, 0.0, self._stream())
                gz = torch.empty_like(dh)
                gz16 = torch.empty(batch, K, dtype=torch.int16, device=_lib.DEVICE)
                _lib.call("c3l_mlp_gelu_backward", _lib.dptr(zs[l - 1]), _lib.dptr(dh), _lib.dptr(gz), _lib.dptr(gz16),
                          batch * K, self._stream())
                g, g16 = gz, gz16
        return gW, gb

    def loss_gradient(self, y, target, batch: int):
        torch = self.torch
        total = y.numel()
        dy = torch.empty_like(y)
        dy16 = torch.empty(y.shape, dtype=torch.int16, device=_lib.DEVICE)
        parts = 64
        loss = torch.empty(parts, dtype=torch.float32, device=_lib.DEVICE)
        _lib.call("c3l_mlp_mse", _lib.dptr(y), _lib.dptr(target), _lib.dptr(dy), _lib.dptr(dy16), _lib.dptr(loss), total,
                  2.0 / total, parts, self._stream())
        return loss, dy, dy16

    def adam(self, gW, gb, lr: float, b1=0.9, b2=0.999, eps=1e-8) -> None:
        self.step_count += 1
        for l in range(len(self.W)):
            _lib.call("c3l_mlp_adam", _lib.dptr(self.W[l]), _lib.dptr(self.mW[l]), _lib.dptr(self.vW[l]), _lib.dptr(gW[l]),
                      _lib.dptr(self.W16[l]), self.W[l].numel(), lr, b1, b2, eps, self.step_count, self._stream())
            _lib.call("c3l_mlp_adam", _lib.dptr(self.b[l]), _lib.dptr(self.mb[l]), _lib.dptr(self.vb[l]), _lib.dptr(gb[l]),
                      None, self.b[l].numel(), lr, b1, b2, eps, self.step_count, self._stream())

    def train_step(self, x16, target, lr: float) -> float:
        batch = x16.shape[0]
        y, cache = self.forward(x16, batch)
        loss, dy, dy16 = self.loss_gradient(y, target, batch)
        gW, gb = self.backward(cache, dy, dy16, batch)
        self.adam(gW, gb, lr)
        return loss

    def to_bf16(self, x32):
        x16 = self.torch.empty(x32.shape, dtype=self.torch.int16, device=_lib.DEVICE)
        _lib.call("c3l_to_bf16", _lib.dptr(x32), _lib.dptr(x16), x32.numel(), self._stream())
        return x16


@dataclass
class WarmStart:
    """Trained map from model parameters to initial real tilde variables."""

    net: AmplitudeMLP
    feature_mean: np.ndarray
    feature_scale: np.ndarray
    mean: object  # torch float64 [D]
    basis: object  # torch float64 [rank, D]
    coefficient_scale: np.ndarray
    train_losses: list
    blend: float = 1.0

    @classmethod
    def fit(cls, problems, amplitudes, *, rank: int = 64, epochs: int = 300, batch: int = 64, lr: float = 1e-3,
            seed: int = 0, validation_fraction: float = 0.25) -> "WarmStart":
        """Train on converged solutions: ``amplitudes[i]`` solve ``problems[i]`` (FactorizedCCProblem).

        The last ``validation_fraction`` of the solutions is held out to fit the blend factor.
        """

        problems, amplitudes = list(problems), list(amplitudes)
        held = int(round(validation_fraction * len(problems)))
        if held <= 0 or held >= len(problems) - 1:
            return cls._train(problems, amplitudes, rank, epochs, batch, lr, seed)
        ws = cls._train(problems[:-held], amplitudes[:-held], rank, epochs, batch, lr, seed)
        num = den = 0.0
        for p, t in zip(problems[-held:], amplitudes[-held:]):
            ws.blend = 1.0
            pred = _canonical_real(p, ws.initial_amplitudes(p))
            truth = _canonical_real(p, t)
            num += float(pred @ truth)
            den += float(pred @ pred)
        ws.blend = float(np.clip(num / den, 0.0, 1.0)) if den > 0.0 else 0.0
        return ws

    @classmethod
    def _train(cls, problems, amplitudes, rank, epochs, batch, lr, seed) -> "WarmStart":
        torch = _lib.require_device()
        X = np.stack([_canonical_real(p, t) for p, t in zip(problems, amplitudes)])
        F = np.stack([model_features(p.model) for p in problems]).astype(np.float64)
        n_samples = X.shape[0]
        if n_samples < 2:
            raise ValueError("at least two training solutions are needed")
        rank = min(rank, n_samples - 1)
        mean = X.mean(axis=0)
        Xc = X - mean
        # thin SVD through the samples' Gram matrix (n_samples << D)
        gram = Xc @ Xc.T
        evals, evecs = np.linalg.eigh(gram)
        order = np.argsort(evals)[::-1][:rank]
        sig = np.sqrt(np.maximum(evals[order], 1e-300))
        basis = (Xc.T @ evecs[:, order]) / sig  # D x rank, orthonormal columns
        coeff = Xc @ basis  # n_samples x rank
        cscale = coeff.std(axis=0) + 1e-12
        fmean = F.mean(axis=0)
        fscale = F.std(axis=0)
        fscale = np.where(fscale > 1e-12, fscale, 1.0)
        Fn = ((F - fmean) / fscale).astype(np.float32)
        Tn = (coeff / cscale).astype(np.float32)
        net = AmplitudeMLP((F.shape[1],) + HIDDEN + (rank,), seed=seed)
        dev = _lib.DEVICE
        Fd = torch.from_numpy(Fn).to(dev)
        Td = torch.from_numpy(Tn).to(dev)
        rng = np.random.default_rng(seed)
        losses = []
        for _ in range(epochs):
            order_ = rng.permutation(n_samples)
            total = 0.0
            for start in range(0, n_samples, batch):
                idx = torch.from_numpy(order_[start:start + batch]).to(dev)
                x16 = net.to_bf16(Fd.index_select(0, idx).contiguous())
                loss = net.train_step(x16, Td.index_select(0, idx).contiguous(), lr)
                total += float(loss.sum().item())
            losses.append(total / (n_samples * rank))
        return cls(net, fmean.astype(np.float32), fscale.astype(np.float32),
                   torch.from_numpy(mean).to(dev), torch.from_numpy(np.ascontiguousarray(basis.T)).to(dev),
                   cscale, losses)

    def coefficients(self, models) -> np.ndarray:
        """Network output (unscaled SVD coefficients) for model parameters."""

        torch = self.net.torch
        F = np.stack([model_features(m) for m in models])
        Fn = ((F - self.feature_mean) / self.feature_scale).astype(np.float32)
        x16 = self.net.to_bf16(torch.from_numpy(Fn).to(_lib.DEVICE))
        y, _ = self.net.forward(x16, len(models))
        return y.cpu().numpy().astype(np.float64) * self.coefficient_scale

    def initial_amplitudes(self, problem) -> np.ndarray:
        """Predicted amplitudes of ``problem`` (FactorizedCCProblem) in its own mode labelling."""

        torch = self.net.torch
        c = torch.from_numpy(self.coefficients([problem.model])[0]).to(_lib.DEVICE)
        r = (self.mean + c @ self.basis).cpu().numpy()
        return self.blend * _from_canonical_real(problem, r)

    def initial_variables(self, problem) -> np.ndarray:
        """Real tilde variables of the predicted amplitudes (input of ``solve_component_ccsd``)."""

        return problem.tilde.pack(self.initial_amplitudes(problem), check=False)

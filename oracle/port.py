"""NumPy executor of lowered contraction plans -- TEST INFRASTRUCTURE / CPU BASELINE ONLY.

Runs exactly the arrays the device library receives (``c3lindblad_amd.plan_arrays``):
padded base scatter, contraction levels with ``np.tensordot``, and the wedge
entries evaluated over all canonical output coordinates at once.  It is the
CPU restatement of the component algorithm of ``src/c3l/cc/component_tensor.py``
timed beside the GPU (``bench.py`` ``cpu_baseline``), and the second checker
of the device kernels at sizes the polynomial oracle cannot reach.

Pins: equals ``oracle.polynomial`` residual/JVP/sigma for n = 1..4 random
inputs (``tests/test_oracle_pins.py``).
"""

from __future__ import annotations

import numpy as np


class PortPlan:
    def __init__(self, lowered):
        self.lp = lowered
        self.n = lowered.n
        self.ws = np.zeros(max(lowered.workspace, 1), dtype=np.complex128)
        for off, vec in zip(lowered.const_offset, lowered.const_vec):
            self.ws[off : off + self.n] = lowered.vectors[vec]

    def load(self, kind: int, v: np.ndarray, *, creation: bool = True) -> None:
        n = self.n
        v = np.asarray(v, dtype=np.complex128)
        for k, off, deg, src, sign in self.lp.bases:
            if k != kind:
                continue
            vals = np.where(src >= 0, v[np.maximum(src, 0)], 0.0) * sign
            self.ws[off : off + src.size] = vals
        for level in self.lp.levels:
            k, out_off, par_off, par_rank, slot, vec, sign = level
            if k != kind:
                continue
            for o, p, r, s, vi, sg in zip(out_off, par_off, par_rank, slot, vec, sign):
                parent = self.ws[p : p + n**r].reshape((n,) * r, order="F")  # first index fastest
                out = np.tensordot(parent, self.lp.vectors[vi], axes=([s], [0]))
                self.ws[o : o + n ** (r - 1)] = sg * np.ravel(out, order="F")
        if kind == 0 and creation:  # amplitudes changed: rebuild the creation tensors that depend on them
            self.apply(4, None)

    def apply_sampled(self, program: int, per_block: int, seed: int = 0):
        """Program values at up to ``per_block`` seeded random coordinates of every output block
        (all coordinates of smaller blocks), without the diagonal term: ``(dest, values)``.

        The sample costs ``per_block`` coordinates per block instead of the whole output, so
        residuals can be spot-checked where the full evaluation would take too long.
        """

        n = self.n
        rng = np.random.default_rng(seed)
        dests, values = [], []
        for pid, deg, coord, dest, coeff, fstart, foff, fcode in self.lp.blocks:
            if pid != program:
                continue
            dest = np.asarray(dest)  # dense blocks list their destinations on demand
            rows = np.arange(dest.size) if dest.size <= per_block else np.sort(
                rng.choice(dest.size, per_block, replace=False))
            sub = coord[rows]
            acc = np.zeros(rows.size, dtype=np.complex128)
            for t in range(coeff.size):
                prod = np.full(rows.size, coeff[t], dtype=np.complex128)
                for f in range(fstart[t], fstart[t + 1]):
                    code = int(fcode[f])
                    flat = np.zeros(rows.size, dtype=np.int64)
                    for slot in range((code & 7) - 1, -1, -1):
                        flat = flat * n + sub[:, (code >> (3 + 3 * slot)) & 7]
                    prod *= self.ws[foff[f] + flat]
                acc += prod
            dests.append(dest[rows])
            values.append(acc)
        return np.concatenate(dests), np.concatenate(values)

    def apply(self, program: int, out: np.ndarray | None, diag=None, x=None) -> np.ndarray:
        """Run one program; ``out=None`` writes a workspace (creation) program in place."""

        n = self.n
        target = self.ws if out is None else out
        for pid, deg, coord, dest, coeff, fstart, foff, fcode in self.lp.blocks:
            if pid != program:
                continue
            if not isinstance(dest, np.ndarray):  # a dense block: consecutive destinations
                dest = slice(dest.start, dest.start + dest.size)
            count = coord.shape[0]
            acc = np.zeros(count, dtype=np.complex128)
            for t in range(coeff.size):
                prod = np.full(count, coeff[t], dtype=np.complex128)
                for f in range(fstart[t], fstart[t + 1]):
                    code = int(fcode[f])
                    rank = code & 7
                    flat = np.zeros(count, dtype=np.int64)
                    for s in range(rank - 1, -1, -1):  # slot s has stride n**s
                        flat = flat * n + coord[:, (code >> (3 + 3 * s)) & 7]
                    prod *= self.ws[foff[f] + flat]
                acc += prod
            if diag is not None:
                acc += diag[dest] * x[dest]
            target[dest] = acc
        return target

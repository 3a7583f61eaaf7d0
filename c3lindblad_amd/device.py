"""Device-side objects: uploaded contraction plans and tilde coordinate maps.

Python/torch own every vector; a ``DevicePlan`` owns the library plan handle
(index maps, coefficients, atom workspace) and frees it on ``close``/GC.
"""

from __future__ import annotations

import ctypes
import itertools

import numpy as np

from . import _lib
from .plan_arrays import KINDS, PROGRAMS, DenseDest, LoweredPlan

_tokens = itertools.count(1)


def _c(array: np.ndarray, dtype) -> np.ndarray:
    return np.ascontiguousarray(array, dtype=dtype)


class DevicePlan:
    """A ``LoweredPlan`` uploaded to the GPU (``c3l_plan_*``)."""

    matrix_cores = 1  # matrix-core mode of the FP32 programs of plans created from now on

    def __init__(self, lowered: LoweredPlan):
        self.torch = _lib.require_device()
        self.lowered = lowered
        self._has_creation = any(b[0] == PROGRAMS["creation"] for b in lowered.blocks)
        self.loaded: dict[int, int] = {}
        handle = ctypes.c_void_p()
        vectors = _c(lowered.vectors, np.complex128)
        const_offset = _c(lowered.const_offset, np.int64)
        const_vec = _c(lowered.const_vec, np.int32)
        _lib.call(
            "c3l_plan_create",
            lowered.n,
            vectors.shape[0],
            _lib.host_ptr(vectors) if vectors.size else None,
            lowered.workspace,
            const_vec.size,
            None,
            _lib.host_ptr(const_offset) if const_offset.size else None,
            _lib.host_ptr(const_vec) if const_vec.size else None,
            ctypes.byref(handle),
        )
        self.handle = handle
        if DevicePlan.matrix_cores != 1:
            self.set_matrix_cores(DevicePlan.matrix_cores)
        for kind, offset, deg, src, sign in lowered.bases:
            s = _c(src, np.int64)
            g = _c(sign, np.int8)
            _lib.call("c3l_plan_add_base", handle, kind, int(offset), deg, _lib.host_ptr(s), _lib.host_ptr(g))
        for kind, out_off, par_off, par_rank, slot, vec, sign in lowered.levels:
            arrays = [
                _c(out_off, np.int64),
                _c(par_off, np.int64),
                _c(par_rank, np.int32),
                _c(slot, np.int32),
                _c(vec, np.int32),
                _c(sign, np.int32),
            ]
            _lib.call("c3l_plan_add_level", handle, kind, arrays[0].size, *[_lib.host_ptr(a) for a in arrays])
        for program, deg, coord, dest, coeff, fstart, foff, fcode in lowered.blocks:
            coeff = _c(coeff, np.complex128)
            fstart = _c(fstart, np.int32)
            foff = _c(foff, np.int64)
            fcode = _c(fcode, np.int32)
            if isinstance(dest, DenseDest):
                _lib.call(
                    "c3l_plan_add_dense_block", handle, program, deg, dest.start, coeff.size,
                    _lib.host_ptr(coeff) if coeff.size else None, _lib.host_ptr(fstart), foff.size,
                    _lib.host_ptr(foff) if foff.size else None, _lib.host_ptr(fcode) if fcode.size else None,
                )
                continue
            coord = _c(coord, np.int32)
            dest = _c(dest, np.int64)
            _lib.call(
                "c3l_plan_add_block",
                handle,
                program,
                deg,
                dest.size,
                _lib.host_ptr(coord) if coord.size else None,
                _lib.host_ptr(dest),
                coeff.size,
                _lib.host_ptr(coeff) if coeff.size else None,
                _lib.host_ptr(fstart),
                foff.size,
                _lib.host_ptr(foff) if foff.size else None,
                _lib.host_ptr(fcode) if fcode.size else None,
            )

    def load(self, kind: str, vector) -> int:
        """Scatter a complex device vector into the ``kind`` bases and build its atoms."""

        _lib.call("c3l_plan_load", self.handle, KINDS[kind], _lib.dptr(vector), _lib.stream_ptr(self.torch))
        if kind == "T" and self.has_creation:
            self.apply("creation", None)  # creation tensors depend on the amplitudes only
        token = next(_tokens)
        self.loaded[KINDS[kind]] = token
        return token

    def apply(self, program: str, out, diag=None, x=None) -> None:
        """Run a program into ``out``; ``out=None`` runs a workspace (creation) program."""

        _lib.call(
            "c3l_plan_apply",
            self.handle,
            PROGRAMS[program],
            None if diag is None else _lib.dptr(diag),
            None if x is None else _lib.dptr(x),
            None if out is None else _lib.dptr(out),
            _lib.stream_ptr(self.torch),
        )

    @property
    def has_creation(self) -> bool:
        return self._has_creation

    def set_precision(self, kind: str, program: str, bits: int) -> None:
        """Build ``kind`` atoms and evaluate ``program`` in FP32 (``bits=32``) or FP64."""

        _lib.call("c3l_plan_set_precision", self.handle, KINDS[kind], PROGRAMS[program], int(bits))
        self.__dict__.setdefault("_bits", {})[program] = int(bits)

    def precision(self, program: str) -> int:
        """Arithmetic bits of ``program`` (64 unless set otherwise)."""

        return self.__dict__.get("_bits", {}).get(program, 64)

    def set_matrix_cores(self, mode) -> None:
        """FP32 programs: rank-4 two-atom blocks on the matrix cores -- ``False``/0 never,
        ``True``/1 blocks large enough to gain (the default), 2 every such block."""

        _lib.call("c3l_plan_set_matrix_cores", self.handle, int(mode))

    def matrix_core_tiles(self, program: str) -> int:
        tiles = ctypes.c_int32()
        _lib.call("c3l_plan_matrix_core_tiles", self.handle, PROGRAMS[program], ctypes.byref(tiles))
        return int(tiles.value)

    def autotune(self, program: str, out, diag=None, x=None, reps: int = 5) -> dict:
        """Choose the launch shape of ``program`` by timing candidates on this plan
        (``c3l_plan_autotune``); the choice applies to every plan with the same digest.
        Returns the default and tuned time per launch (ms) and the chosen shape."""

        ms_default, ms_best = ctypes.c_double(), ctypes.c_double()
        shape = (ctypes.c_int32 * 6)()
        _lib.call(
            "c3l_plan_autotune", self.handle, PROGRAMS[program], None if diag is None else _lib.dptr(diag),
            None if x is None else _lib.dptr(x), _lib.dptr(out), int(reps), _lib.stream_ptr(self.torch),
            ctypes.byref(ms_default), ctypes.byref(ms_best), shape,
        )
        names = ("lanes_max_log2", "lane_split_workgroups", "coordinates_per_thread", "coordinate_workgroups",
                 "xcd_slicing", "matrix_core_min_entries")
        return {"program": program, "ms_default": ms_default.value, "ms_tuned": ms_best.value,
                "shape": dict(zip(names, (int(v) for v in shape)))}

    def set_timing(self, enable: bool) -> None:
        _lib.call("c3l_plan_set_timing", self.handle, 1 if enable else 0)

    def block_stats(self) -> list[dict]:
        """Per output block: program, degree, coordinates, entries, factors."""

        lib = _lib.load_library()
        names = {v: k for k, v in PROGRAMS.items()}
        out = []
        for i in range(lib.c3l_plan_block_count(self.handle)):
            program, deg, n_entries, n_factors = (ctypes.c_int32() for _ in range(4))
            count = ctypes.c_int64()
            _lib.call(
                "c3l_plan_block_info", self.handle, i, ctypes.byref(program), ctypes.byref(deg), ctypes.byref(count),
                ctypes.byref(n_entries), ctypes.byref(n_factors),
            )
            out.append(dict(program=names[program.value], deg=deg.value, count=count.value, entries=n_entries.value,
                            factors=n_factors.value))
        return out

    def program_time(self, program: str) -> tuple[float, int]:
        """Accumulated kernel milliseconds and launches of one program since timing was enabled."""

        ms, launches = ctypes.c_double(), ctypes.c_int64()
        _lib.call("c3l_plan_program_time", self.handle, PROGRAMS[program], ctypes.byref(ms), ctypes.byref(launches))
        return ms.value, launches.value

    def close(self) -> None:
        if getattr(self, "handle", None) is not None and self.handle.value:
            _lib.load_library().c3l_plan_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceTilde:
    """Tilde orbit maps on the device: real coordinates <-> complex amplitudes."""

    def __init__(self, tilde):
        torch = _lib.require_device()
        self.torch = torch
        dev = torch.device(_lib.DEVICE)
        self.dim = tilde.dimension
        self.real_dim = tilde.real_size
        ab = np.stack([tilde.alpha.real, tilde.alpha.imag, tilde.beta.real, tilde.beta.imag], axis=1)
        self.ab = torch.from_numpy(np.ascontiguousarray(ab)).to(dev)
        self.k0 = torch.from_numpy(tilde.k0.astype(np.int64)).to(dev)
        self.k1 = torch.from_numpy(tilde.k1.astype(np.int64)).to(dev)
        self.src = torch.from_numpy(tilde.pack_src.astype(np.int64)).to(dev)
        self.w = torch.from_numpy(np.ascontiguousarray(tilde.pack_w)).to(dev)

    def unpack(self, x, t) -> None:
        _lib.call(
            "c3l_tilde_unpack",
            _lib.dptr(x),
            _lib.dptr(t),
            _lib.dptr(self.k0),
            _lib.dptr(self.k1),
            _lib.dptr(self.ab),
            self.dim,
            _lib.stream_ptr(self.torch),
        )

    def precondition(self, x, y, inv_diag) -> None:
        """``y = pack(unpack(x) * inv_diag)`` in one kernel."""

        _lib.call(
            "c3l_tilde_precondition",
            _lib.dptr(x),
            _lib.dptr(y),
            _lib.dptr(self.k0),
            _lib.dptr(self.k1),
            _lib.dptr(self.ab),
            _lib.dptr(self.src),
            _lib.dptr(self.w),
            _lib.dptr(inv_diag),
            self.real_dim,
            _lib.stream_ptr(self.torch),
        )

    def pack(self, t, x) -> None:
        _lib.call(
            "c3l_tilde_pack",
            _lib.dptr(t),
            _lib.dptr(x),
            _lib.dptr(self.src),
            _lib.dptr(self.w),
            self.real_dim,
            _lib.stream_ptr(self.torch),
        )


class KernelTimer:
    """Optional HIP-event timing of Krylov kernels on the launching (current) stream."""

    enabled = False
    _pending: list = []

    @classmethod
    def start(cls):
        if not cls.enabled:
            return None
        import torch

        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        return ev

    @classmethod
    def stop(cls, name: str, start, nbytes: float) -> None:
        if start is None:
            return
        import torch

        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        cls._pending.append((name, start, ev, nbytes))

    @classmethod
    def summary(cls) -> dict:
        """``{name: {"ms": total, "launches": k, "bytes": total algorithmic bytes}}``."""

        out: dict = {}
        for name, a, b, nbytes in cls._pending:
            b.synchronize()
            d = out.setdefault(name, {"ms": 0.0, "launches": 0, "bytes": 0.0})
            d["ms"] += a.elapsed_time(b)
            d["launches"] += 1
            d["bytes"] += nbytes
        cls._pending = []
        return out


class Reducer:
    """Scratch-backed multi-dot and multi-axpy over contiguous Krylov bases."""

    def __init__(self, dim: int, max_cols: int, complex_: bool):
        torch = _lib.require_device()
        self.torch = torch
        self.dim = int(dim)
        self.complex = complex_
        chunks = int(_lib.load_library().c3l_reduction_chunks(self.dim))
        dtype = torch.complex128 if complex_ else torch.float64
        self.scratch = torch.empty(chunks * max_cols, dtype=dtype, device=_lib.DEVICE)
        self.coef = torch.empty(max_cols, dtype=dtype, device=_lib.DEVICE)

    def dot(self, V, ncols: int, w, out, conjugate: bool = True) -> None:
        ld = V.stride(0)
        ev = KernelTimer.start()
        if self.complex:
            _lib.call(
                "c3l_zmultidot", _lib.dptr_rows(V), ld, ncols, _lib.dptr(w), self.dim, 1 if conjugate else 0,
                _lib.dptr(self.scratch), _lib.dptr(out), _lib.stream_ptr(self.torch),
            )
        else:
            _lib.call(
                "c3l_rmultidot", _lib.dptr_rows(V), ld, ncols, _lib.dptr(w), self.dim,
                _lib.dptr(self.scratch), _lib.dptr(out), _lib.stream_ptr(self.torch),
            )
        elem = 16 if self.complex else 8
        KernelTimer.stop("zmultidot" if self.complex else "rmultidot", ev, float((ncols + 1) * self.dim * elem))

    def axpy(self, V, ncols: int, alpha, w, scale: float = 1.0) -> None:
        """``w += scale * sum_k alpha[k] V[k]``."""

        ld = V.stride(0)
        name = "c3l_zmultiaxpy" if self.complex else "c3l_rmultiaxpy"
        ev = KernelTimer.start()
        _lib.call(name, _lib.dptr(V), ld, ncols, _lib.dptr(alpha), float(scale), _lib.dptr(w), self.dim,
                  _lib.stream_ptr(self.torch))
        elem = 16 if self.complex else 8
        KernelTimer.stop(name[4:], ev, float((ncols + 2) * self.dim * elem))


def rscale(x, y, a: float, torch) -> None:
    _lib.call("c3l_rscale", _lib.dptr(x), _lib.dptr(y), float(a), x.numel(), _lib.stream_ptr(torch))


def raxpby(a: float, x, b: float, y, torch) -> None:
    _lib.call("c3l_raxpby", float(a), _lib.dptr(x), float(b), _lib.dptr(y), x.numel(), _lib.stream_ptr(torch))

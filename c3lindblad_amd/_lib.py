"""ctypes binding of ``libc3l_hip.so`` (declared in ``include/c3l_hip.h``).

The library is loaded from the package directory.  There is no fallback: if
it is missing or no GPU is visible, every device entry point raises.
Status codes map to the reference's exception types: invalid shapes or
indices -> ``ValueError``, allocation failure -> ``MemoryError``, HIP errors
-> ``RuntimeError``.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

LIB_NAME = "libc3l_hip.so"
LIB_PATH = Path(__file__).resolve().parent / LIB_NAME

C3L_OK, C3L_EINVAL, C3L_EHIP, C3L_ENOMEM = 0, 1, 2, 3

_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64

SIGNATURES = {
    "c3l_last_error": ([], ctypes.c_char_p),
    "c3l_version": ([], ctypes.c_char_p),
    "c3l_plan_create": ([_I32, _I32, _P, _I64, _I32, _P, _P, _P, ctypes.POINTER(_P)], ctypes.c_int),
    "c3l_plan_destroy": ([_P], ctypes.c_int),
    "c3l_plan_add_base": ([_P, _I32, _I64, _I32, _P, _P], ctypes.c_int),
    "c3l_plan_add_level": ([_P, _I32, _I32, _P, _P, _P, _P, _P, _P], ctypes.c_int),
    "c3l_plan_add_block": ([_P, _I32, _I32, _I64, _P, _P, _I32, _P, _P, _I32, _P, _P], ctypes.c_int),
    "c3l_plan_add_dense_block": ([_P, _I32, _I32, _I64, _I32, _P, _P, _I32, _P, _P], ctypes.c_int),
    "c3l_plan_load": ([_P, _I32, _P, _P], ctypes.c_int),
    "c3l_plan_apply": ([_P, _I32, _P, _P, _P, _P], ctypes.c_int),
    "c3l_plan_set_timing": ([_P, _I32], ctypes.c_int),
    "c3l_plan_set_precision": ([_P, _I32, _I32, _I32], ctypes.c_int),
    "c3l_plan_set_matrix_cores": ([_P, _I32], ctypes.c_int),
    "c3l_plan_matrix_core_tiles": ([_P, _I32, _P], ctypes.c_int),
    "c3l_plan_autotune": ([_P, _I32, _P, _P, _P, _I32, _P, _P, _P, _P], ctypes.c_int),
    "c3l_plan_block_count": ([_P], _I32),
    "c3l_plan_block_info": ([_P, _I32, _P, _P, _P, _P, _P], ctypes.c_int),
    "c3l_plan_program_time": ([_P, _I32, _P, _P], ctypes.c_int),
    "c3l_tilde_unpack": ([_P, _P, _P, _P, _P, _I64, _P], ctypes.c_int),
    "c3l_tilde_pack": ([_P, _P, _P, _P, _I64, _P], ctypes.c_int),
    "c3l_tilde_pack_precondition": ([_P, _P, _P, _P, _P, _I64, _P], ctypes.c_int),
    "c3l_rmultidot": ([_P, _I64, _I32, _P, _I64, _P, _P, _P], ctypes.c_int),
    "c3l_rmultiaxpy": ([_P, _I64, _I32, _P, ctypes.c_double, _P, _I64, _P], ctypes.c_int),
    "c3l_zmultidot": ([_P, _I64, _I32, _P, _I64, _I32, _P, _P, _P], ctypes.c_int),
    "c3l_zmultidot_noalloc": ([_P, _I64, _I32, _P, _I64, _I32, _P, _P], ctypes.c_int),
    "c3l_zmultiaxpy": ([_P, _I64, _I32, _P, ctypes.c_double, _P, _I64, _P], ctypes.c_int),
    "c3l_tilde_precondition": ([_P, _P, _P, _P, _P, _P, _P, _P, _I64, _P], ctypes.c_int),
    "c3l_rscale": ([_P, _P, ctypes.c_double, _I64, _P], ctypes.c_int),
    "c3l_raxpby": ([ctypes.c_double, _P, ctypes.c_double, _P, _I64, _P], ctypes.c_int),
    "c3l_reduction_chunks": ([_I64], _I64),
    "c3l_shifted_triangular_green": ([_P, _I32, _P, _P, _P, _P, _I32, _P, _P, _P], ctypes.c_int),
    "c3l_gmres": (
        [_P, _P, _P, _P, _P, _P, _P, _P, _I32, _P, _P, _P, _P, ctypes.c_double, ctypes.c_double, _I32, _I32, _P, _P, _P,
         _P, _P],
        ctypes.c_int,
    ),
    "c3l_gmres_group": (
        [_P, _P, _I32, _P, _P, _P, _P, _P, _P, _I32, _P, _P, _P, _P, _P, _P, _P, _I32, _I32, _P, _P, _P, _P, _P],
        ctypes.c_int,
    ),
    "c3l_bf16_gemm": ([_P, _I64, _I64, _P, _I64, _I64, _P, _I64, _I32, _I32, _I32, ctypes.c_float, ctypes.c_float, _P],
                      ctypes.c_int),
    "c3l_mlp_bias_act": ([_P, _P, _P, _I32, _I32, _I32, _P], ctypes.c_int),
    "c3l_mlp_gelu_backward": ([_P, _P, _P, _P, _I64, _P], ctypes.c_int),
    "c3l_mlp_column_sum": ([_P, _P, _I32, _I32, _P], ctypes.c_int),
    "c3l_mlp_mse": ([_P, _P, _P, _P, _P, _I64, ctypes.c_float, _I32, _P], ctypes.c_int),
    "c3l_mlp_adam": ([_P, _P, _P, _P, _P, _I64, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, _I32, _P],
                     ctypes.c_int),
    "c3l_to_bf16": ([_P, _P, _I64, _P], ctypes.c_int),
    "c3l_krylov_workspace": ([_I32, _I64, _I32, _I32, _P, _P], ctypes.c_int),
    "c3l_krylov_basis_ld": ([_I64], _I64),
    "c3l_krylov_set_timing": ([_I32], ctypes.c_int),
    "c3l_krylov_timing": ([_I32, _P, _P, _P], ctypes.c_int),
    "c3l_arnoldi": (
        [_P, _P, _P, _P, _I64, _I32, _P, _I32, _P, ctypes.c_double, _P, _P, _P, _P, _P, _P, _P, _P, _P,
         ctypes.c_double, _I32, ctypes.c_double, _I32, _P, _P, _P, _P],
        ctypes.c_int,
    ),
    "c3l_arnoldi_batch": (
        [_P, _I32, _P, _P, _P, _I64, _I32, _P, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
         ctypes.c_double, _I32, ctypes.c_double, _P, _P, _P, _P],
        ctypes.c_int,
    ),
    "c3l_arnoldi_begin": (
        [_P, _I32, _P, _P, _P, _I64, _I32, _P, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
         ctypes.c_double, _I32, ctypes.c_double, _P, _P],
        ctypes.c_int,
    ),
    "c3l_arnoldi_group_begin": (
        [_P, _I32, _I32, _P, _P, _P, _P, _I64, _I32, _P, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
         ctypes.c_double, _I32, ctypes.c_double, _P, _P],
        ctypes.c_int,
    ),
    "c3l_arnoldi_drive": ([_P, _I32], ctypes.c_int),
    "c3l_arnoldi_end": ([_P, _P, _P, _P], ctypes.c_int),
}


class TildeMaps(ctypes.Structure):
    """``c3l_tilde_maps`` of ``include/c3l_hip.h``."""

    _fields_ = [
        ("dim", ctypes.c_int64),
        ("real_dim", ctypes.c_int64),
        ("k0", ctypes.c_void_p),
        ("k1", ctypes.c_void_p),
        ("ab", ctypes.c_void_p),
        ("src", ctypes.c_void_p),
        ("w", ctypes.c_void_p),
        ("inv_diag", ctypes.c_void_p),
    ]

_lib = None

# Placement of every device tensor of the package.
DEVICE = "cuda"


def pinned(tensor):
    """Page-locked host staging buffer for device-to-host scalar transfers."""

    return tensor.pin_memory() if DEVICE == "cuda" else tensor


def load_library() -> ctypes.CDLL:
    """Load and type the shared library (no GPU needed to load)."""

    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build the HIP extension first (python -c 'import __graft_entry__ as g; g.build()')"
            )
        lib = ctypes.CDLL(str(LIB_PATH))
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
    return _lib


def check(status: int) -> None:
    if status == C3L_OK:
        return
    message = load_library().c3l_last_error().decode()
    if status == C3L_EINVAL:
        raise ValueError(message)
    if status == C3L_ENOMEM:
        raise MemoryError(message)
    raise RuntimeError(message)


def call(name: str, *args) -> None:
    check(getattr(load_library(), name)(*args))


def host_ptr(array: np.ndarray) -> int:
    if not array.flags.c_contiguous:
        raise ValueError("host array must be contiguous")
    return array.ctypes.data


def krylov_basis_ld(rows: int) -> int:
    """Leading dimension (entries) of a c3l_gmres basis of ``rows`` real coordinates."""

    return int(load_library().c3l_krylov_basis_ld(int(rows)))


def krylov_workspace(complex_basis: bool, dim: int, max_columns: int, n_shifts: int = 0) -> tuple[int, int]:
    """``(scratch, coef)`` sizes in doubles of ``c3l_arnoldi`` / ``c3l_gmres`` workspaces."""

    scratch, coef = ctypes.c_int64(), ctypes.c_int64()
    call("c3l_krylov_workspace", 1 if complex_basis else 0, int(dim), int(max_columns), int(n_shifts),
         ctypes.byref(scratch), ctypes.byref(coef))
    return scratch.value, coef.value


def require_device():
    """Return torch after checking that a GPU is visible (the product path has no CPU fallback)."""

    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("the c3lindblad_amd device path needs a visible MI355X GPU (no CPU fallback)")
    load_library()
    return torch


def dptr(tensor) -> int:
    if not tensor.is_cuda or not tensor.is_contiguous():
        raise ValueError("device tensor must be a contiguous GPU tensor")
    return tensor.data_ptr()


def dptr_rows(tensor) -> int:
    """Pointer of a 2-D GPU tensor whose rows are contiguous (any row stride, passed separately as
    the leading dimension), e.g. a basis view of padded columns."""

    if not tensor.is_cuda or tensor.dim() != 2 or tensor.stride(1) != 1 or tensor.stride(0) < tensor.shape[1]:
        raise ValueError("device tensor must be a GPU matrix with contiguous rows")
    return tensor.data_ptr()


def stream_ptr(torch) -> int:
    return torch.cuda.current_stream().cuda_stream


def env_flag(name: str) -> bool:
    return os.environ.get(name, "") not in ("", "0")

"""Checkpointed, resumable Arnoldi resolvent with the basis mirrored to disk.

Drop-in for ``checkpointed_arnoldi_resolvent`` (``src/c3l/green/checkpointed_arnoldi.py:168-464``)
and ``eom_checkpoint_identity`` (``:101-131``): the same arguments, validation, files
(``basis.npy`` memory map of shape ``(n, m+1)``, column major; ``state.npz``), binding of a
checkpoint to its inputs, ``maximum_new_steps`` interruption and residual checks at the declared
``checkpoint_dimensions`` (Galerkin residuals from the Schur form of ``H_m``).

Every step runs on the GPU: the sigma action (``matvec_device``), two classical Gram--Schmidt
passes (``c3l_zmultidot`` / ``c3l_zmultiaxpy``), the norm and the sink projections.  The new
basis vector is copied to the memory map after each step; a resumed run uploads the saved
columns and continues with the same kernels on the same bits, so an interrupted and resumed
run equals an uninterrupted one exactly.
"""

from __future__ import annotations

import hashlib
import os
import resource
from dataclasses import dataclass
from pathlib import Path
from time import perf_counter

import numpy as np
import scipy.linalg

from . import _lib
from .device import Reducer
from .response import ArnoldiResolvent

ORTHOGONALIZATION_BACKEND = "hip_cgs2"


def require_resources(*, required_memory_bytes: int, required_disk_bytes: int, available_memory_bytes: int,
                      free_disk_bytes: int) -> None:
    """Reject a run when either declared resource is one byte short (``checkpointed_arnoldi.py:67-85``)."""

    if available_memory_bytes < required_memory_bytes:
        raise MemoryError(f"insufficient memory: need {required_memory_bytes}, have {available_memory_bytes} bytes")
    if free_disk_bytes < required_disk_bytes:
        raise OSError(f"insufficient disk: need {required_disk_bytes}, have {free_disk_bytes} bytes")


def _hash_array(digest, array) -> None:
    value = np.ascontiguousarray(array)
    digest.update(repr((value.shape, value.dtype.str)).encode())
    digest.update(value.view(np.uint8))


def eom_checkpoint_identity(context) -> str:
    """SHA-256 of the inputs that define an EOM context's sigma action: rapidities, amplitudes,
    product coefficients and linear forms, and the response coordinates."""

    digest = hashlib.sha256(b"c3lindblad_amd-eom-context")
    cc = context.problem.static.oracle
    _hash_array(digest, np.asarray(cc.rapidities))
    _hash_array(digest, np.asarray(context.amplitudes))
    for record in cc.products:
        digest.update(repr((complex(record.coefficient), tuple(f.components for f in record.factors))).encode())
    coords = context.problem.coordinates
    digest.update(repr((coords.n, coords.dimension, context.problem.charge)).encode())
    for block in coords.blocks:
        digest.update(repr(block.key).encode())
        _hash_array(digest, block.positions)
        _hash_array(digest, block.indices)
    return digest.hexdigest()


def _binding(b, sinks, z, *, n, maximum_dimension, checkpoint_dimensions, residual_tolerance, breakdown_tolerance,
             provenance) -> str:
    digest = hashlib.sha256()
    for value in (b, sinks, z):
        _hash_array(digest, value)
    digest.update(repr((n, maximum_dimension, checkpoint_dimensions, residual_tolerance, breakdown_tolerance,
                        provenance)).encode())
    return digest.hexdigest()


def _write_state(path: Path, **arrays) -> None:
    temporary = path.with_name(path.name + ".tmp.npz")
    with temporary.open("wb") as stream:
        np.savez(stream, **arrays)
    os.replace(temporary, path)


def _reduced_residuals(H: np.ndarray, beta: float, shifts: np.ndarray) -> np.ndarray:
    """Galerkin residuals ``|h_{m+1,m} e_m^T (z - H_m)^{-1} beta e_1| / beta`` (``green/krylov.py:257-281``)."""

    m = H.shape[1]
    T, Q = scipy.linalg.schur(H[:m, :m], output="complex", check_finite=False)
    rhs = Q.conj().T[:, 0] * beta
    identity = np.eye(m, dtype=np.complex128)
    out = np.empty(shifts.size)
    for p, shift in enumerate(shifts):
        y = Q @ scipy.linalg.solve_triangular(shift * identity - T, rhs, lower=False, check_finite=False)
        out[p] = abs(H[m, m - 1] * y[-1]) / max(beta, np.finfo(np.float64).tiny)
    return out


@dataclass
class CheckpointedArnoldiResult:
    """``checkpointed_arnoldi.py:46-64``: the projection, every sink's ``sink . v_k`` and timings."""

    projection: ArnoldiResolvent
    sink_projections: np.ndarray  # (sinks, m)
    orthogonalization_wall_seconds: np.ndarray
    checkpoint_dimensions: tuple
    directory: Path
    orthogonalization_backend: str = ORTHOGONALIZATION_BACKEND

    def transfer(self, shifts, *, dimension: int | None = None) -> np.ndarray:
        m = self.projection._dimension(dimension)
        y = self.projection.coefficients(shifts, dimension=m)
        return np.asarray(y @ self.sink_projections[:, :m].T, dtype=np.complex128)


def checkpointed_arnoldi_resolvent(
    operator,
    rhs,
    sinks,
    shifts,
    directory,
    *,
    maximum_dimension: int,
    checkpoint_dimensions,
    residual_tolerance: float,
    breakdown_tolerance: float = 1.0e-14,
    maximum_new_steps: int | None = None,
    provenance: str = "",
) -> CheckpointedArnoldiResult:
    """Build ``K_m(sigma, b)`` on the GPU with a disk mirror of the basis, resuming the last checkpoint
    in ``directory`` when one exists for the same inputs (``checkpointed_arnoldi.py:168-464``)."""

    torch = _lib.require_device()
    if not hasattr(operator, "dimension") or not hasattr(operator, "matvec_device"):
        raise TypeError("operator must expose dimension and a device matvec")
    n = int(operator.dimension)
    b = np.asarray(rhs, dtype=np.complex128)
    rows = np.asarray(sinks, dtype=np.complex128)
    z = np.asarray(shifts, dtype=np.complex128)
    if b.shape != (n,) or not np.all(np.isfinite(b)):
        raise ValueError("rhs must be a finite vector matching the operator")
    if rows.ndim != 2 or rows.shape[1] != n or not np.all(np.isfinite(rows)):
        raise ValueError("sinks must be a finite matrix with one row per sink")
    if z.ndim != 1 or z.size == 0 or not np.all(np.isfinite(z)):
        raise ValueError("shifts must be a finite nonempty vector")
    if isinstance(maximum_dimension, bool) or not isinstance(maximum_dimension, int) or not 1 <= maximum_dimension <= n:
        raise ValueError("maximum_dimension must lie in [1, operator.dimension]")
    checks = tuple(checkpoint_dimensions)
    if (not checks or checks != tuple(sorted(set(checks)))
            or any(isinstance(m, bool) or not isinstance(m, int) for m in checks)
            or checks[0] < 1 or checks[-1] > maximum_dimension):
        raise ValueError("checkpoint_dimensions must be sorted unique valid prefixes")
    tolerance = float(residual_tolerance)
    breakdown_threshold = float(breakdown_tolerance)
    if not np.isfinite(tolerance) or tolerance <= 0.0:
        raise ValueError("residual_tolerance must be finite and positive")
    if not np.isfinite(breakdown_threshold) or breakdown_threshold <= 0.0:
        raise ValueError("breakdown_tolerance must be finite and positive")
    if maximum_new_steps is not None and (isinstance(maximum_new_steps, bool) or not isinstance(maximum_new_steps, int)
                                          or maximum_new_steps < 0):
        raise ValueError("maximum_new_steps must be a nonnegative integer")
    beta = float(np.linalg.norm(b))
    if beta == 0.0:
        raise ValueError("rhs must be nonzero")

    root = Path(directory)
    root.mkdir(parents=True, exist_ok=True)
    basis_path = root / "basis.npy"
    state_path = root / "state.npz"
    key = _binding(b, rows, z, n=n, maximum_dimension=maximum_dimension, checkpoint_dimensions=checks,
                   residual_tolerance=tolerance, breakdown_tolerance=breakdown_threshold, provenance=provenance)
    m_max = maximum_dimension
    H = np.zeros((m_max + 1, m_max), dtype=np.complex128)
    projected = np.zeros((rows.shape[0], m_max + 1), dtype=np.complex128)
    ortho_seconds = np.zeros(m_max)
    check_m: list[int] = []
    check_r: list[float] = []
    elapsed = 0.0
    breakdown = False
    dev = _lib.DEVICE
    V = torch.empty(m_max + 1, n, dtype=torch.complex128, device=dev)  # rows are basis vectors
    if state_path.exists():
        if not basis_path.exists():
            raise ValueError("checkpoint basis is missing")
        state = np.load(state_path, allow_pickle=False)
        if str(state["binding"].item()) != key:
            raise ValueError("checkpoint inputs do not match the requested problem")
        if str(state["orthogonalization_backend"].item()) != ORTHOGONALIZATION_BACKEND:
            raise ValueError("checkpoint orthogonalization backend does not match")
        actual = int(state["dimension"].item())
        H[: actual + 1, :actual] = state["hessenberg"]
        projected[:, : actual + 1] = state["sink_projections"]
        ortho_seconds[:actual] = state["orthogonalization_wall_seconds"]
        check_m = [int(v) for v in state["check_dimensions"]]
        check_r = [float(v) for v in state["check_residuals"]]
        elapsed = float(state["construction_wall_seconds"].item())
        breakdown = bool(state["breakdown"].item())
        mirror = np.lib.format.open_memmap(basis_path, mode="r+")
        if mirror.shape != (n, m_max + 1) or mirror.dtype != np.complex128:
            raise ValueError("checkpoint basis shape or dtype does not match")
        V[: actual + 1].copy_(torch.from_numpy(np.ascontiguousarray(mirror[:, : actual + 1].T)))
    else:
        if basis_path.exists():
            raise ValueError("checkpoint state is missing")
        if maximum_new_steps == 0:
            raise ValueError("zero new steps requires an existing checkpoint")
        actual = 0
        mirror = np.lib.format.open_memmap(basis_path, mode="w+", dtype=np.complex128, shape=(n, m_max + 1),
                                           fortran_order=True)
        V[0].copy_(torch.from_numpy(b / beta))
        mirror[:, 0] = V[0].cpu().numpy()
        mirror.flush()
    sinks_d = torch.from_numpy(np.ascontiguousarray(rows)).to(dev)
    red = Reducer(n, m_max + 1, complex_=True)
    sink_red = Reducer(n, max(rows.shape[0], 1), complex_=True)
    coef = torch.empty(m_max + 1, dtype=torch.complex128, device=dev)
    norm2 = torch.empty(1, dtype=torch.complex128, device=dev)
    w = torch.empty(n, dtype=torch.complex128, device=dev)
    sink_v = torch.empty(max(rows.shape[0], 1), dtype=torch.complex128, device=dev)
    if actual == 0 and rows.shape[0]:
        sink_red.dot(sinks_d, rows.shape[0], V[0], sink_v, conjugate=False)
        projected[:, 0] = sink_v[: rows.shape[0]].cpu().numpy()
    call_started = perf_counter()
    start = actual

    def save() -> None:
        mirror.flush()
        _write_state(
            state_path,
            binding=np.asarray(key),
            dimension=np.asarray(actual, dtype=np.int64),
            hessenberg=H[: actual + 1, :actual],
            sink_projections=projected[:, : actual + 1],
            orthogonalization_wall_seconds=ortho_seconds[:actual],
            check_dimensions=np.asarray(check_m, dtype=np.int64),
            check_residuals=np.asarray(check_r, dtype=np.float64),
            construction_wall_seconds=np.asarray(elapsed + perf_counter() - call_started, dtype=np.float64),
            breakdown=np.asarray(breakdown),
            orthogonalization_backend=np.asarray(ORTHOGONALIZATION_BACKEND),
        )

    while actual < m_max and not breakdown and (maximum_new_steps is None or actual - start < maximum_new_steps):
        j = actual
        operator.matvec_device(V[j], w)
        t0 = perf_counter()
        for _ in range(2):  # classical Gram-Schmidt, two passes (krylov.py:361-370)
            red.dot(V[: j + 1], j + 1, w, coef[: j + 1], conjugate=True)
            red.axpy(V[: j + 1], j + 1, coef[: j + 1], w, scale=-1.0)
            H[: j + 1, j] += coef[: j + 1].cpu().numpy()
        red.dot(w.view(1, -1), 1, w, norm2, conjugate=True)
        h = float(np.sqrt(max(norm2.cpu().numpy()[0].real, 0.0)))
        ortho_seconds[j] = perf_counter() - t0
        if not np.isfinite(h):
            raise ValueError("operator matvec returned an invalid vector")
        H[j + 1, j] = h
        actual = j + 1
        if h <= breakdown_threshold:
            breakdown = True
        else:
            torch.view_as_real(V[actual]).copy_(torch.view_as_real(w) / h)
            if rows.shape[0]:
                sink_red.dot(sinks_d, rows.shape[0], V[actual], sink_v, conjugate=False)
                projected[:, actual] = sink_v[: rows.shape[0]].cpu().numpy()
            mirror[:, actual] = V[actual].cpu().numpy()
        stop_for_limit = maximum_new_steps is not None and actual - start >= maximum_new_steps
        if actual in checks or actual == m_max or breakdown or stop_for_limit:
            residual = float(np.max(_reduced_residuals(H[: actual + 1, :actual], beta, z)))
            if check_m and check_m[-1] == actual:
                check_r[-1] = residual
            else:
                check_m.append(actual)
                check_r.append(residual)
            save()
            if residual <= tolerance or stop_for_limit:
                break
    if not check_m or check_m[-1] != actual:
        residual = float(np.max(_reduced_residuals(H[: actual + 1, :actual], beta, z)))
        check_m.append(actual)
        check_r.append(residual)
        save()
    projection = ArnoldiResolvent(
        basis=V[: actual + 1],
        hessenberg=np.ascontiguousarray(H[: actual + 1, :actual]),
        source_norm=beta,
        breakdown=breakdown,
        shifts=z,
        green=np.full(z.size, np.nan + 0j),
        shift_relative_residuals=_reduced_residuals(H[: actual + 1, :actual], beta, z),
        converged=bool(check_r[-1] <= tolerance),
        construction_wall_seconds=elapsed + perf_counter() - call_started,
        action_calls=actual - start,
        residual_tolerance=tolerance,
        criterion="galerkin",
        peak_rss_kib=int(resource.getrusage(resource.RUSAGE_SELF).ru_maxrss),
    )
    projection.__dict__["_checks"] = (tuple(check_m), np.asarray(check_r))
    return CheckpointedArnoldiResult(
        projection=projection,
        sink_projections=np.asarray(projected[:, :actual]),
        orthogonalization_wall_seconds=np.asarray(ortho_seconds[:actual]),
        checkpoint_dimensions=checks,
        directory=root,
    )

"""Allocation-free complex multidot (``c3l_zmultidot_noalloc``) against ``c3l_zmultidot``.

Both run the same kernels; the allocation-free entry point keeps its partial sums in the
library's own device array, so its results must equal the scratch-buffer entry point's bit for
bit, on consecutive calls that reuse the array with a smaller second input, and sizes beyond the
array are refused before anything is launched.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(torch, ncols, dim, ld, conjugate, seed, dense=True):
    from c3lindblad_amd import _lib

    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    V = torch.randn(ncols, ld, dtype=torch.complex128, device="cuda", generator=g)
    w = torch.randn(dim, dtype=torch.complex128, device="cuda", generator=g)
    chunks = int(_lib.load_library().c3l_reduction_chunks(dim))
    scratch = torch.empty(2 * ncols * chunks, dtype=torch.float64, device="cuda")
    want = torch.empty(ncols, dtype=torch.complex128, device="cuda")
    got = torch.full((ncols,), complex(np.nan, np.nan), dtype=torch.complex128, device="cuda")
    stream = _lib.stream_ptr(torch)
    _lib.call("c3l_zmultidot", _lib.dptr(V), ld, ncols, _lib.dptr(w), dim, conjugate, _lib.dptr(scratch),
              _lib.dptr(want), stream)
    _lib.call("c3l_zmultidot_noalloc", _lib.dptr(V), ld, ncols, _lib.dptr(w), dim, conjugate, _lib.dptr(got), stream)
    torch.cuda.synchronize()
    reference = None
    if dense:
        Vh = V.cpu().numpy()[:, :dim]
        reference = (Vh.conj() if conjugate else Vh) @ w.cpu().numpy()
    return got.cpu().numpy(), want.cpu().numpy(), reference


def test_noalloc_multidot_equals_the_scratch_buffer_entry_point(torch_gpu):
    torch = torch_gpu
    print("case noalloc multidot: 37 x 23150 then 5 x 1000", flush=True)
    for ncols, dim, ld, conjugate, seed in ((37, 23150, 23157, 1, 1), (5, 1000, 1000, 0, 2), (300, 4099, 4100, 1, 3)):
        got, want, dense = _pair(torch, ncols, dim, ld, conjugate, seed)
        assert np.array_equal(got, want), (ncols, dim)
        assert np.max(np.abs(got - dense)) <= 1e-12 * np.max(np.abs(dense))


def test_noalloc_multidot_refuses_sizes_beyond_its_array(torch_gpu):
    """600 columns of 1,789,952 rows need 600 * 874 complex partials (> 2^19): refused with
    ValueError before any launch (the buffers are real and large enough, so nothing could be
    read out of bounds even if a kernel ran)."""

    from c3lindblad_amd import _lib

    torch = torch_gpu
    print("case noalloc multidot: refused size", flush=True)
    ncols, dim = 600, 874 * 2048
    V = torch.empty(ncols, dim, dtype=torch.complex128, device="cuda")
    w = torch.zeros(dim, dtype=torch.complex128, device="cuda")
    out = torch.zeros(ncols, dtype=torch.complex128, device="cuda")
    with pytest.raises(ValueError, match="scratch array"):
        _lib.call("c3l_zmultidot_noalloc", _lib.dptr(V), dim, ncols, _lib.dptr(w), dim, 1, _lib.dptr(out),
                  _lib.stream_ptr(torch))
    del V
    torch.cuda.empty_cache()
    # the largest accepted size for 600 columns still works
    got, want, _ = _pair(torch, 600, 873 * 2048, 873 * 2048, 0, 4, dense=False)
    assert np.array_equal(got, want)

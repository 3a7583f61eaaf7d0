"""The C-ABI library loads without a GPU and exports every declared entry point."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols() -> list[str]:
    text = (ROOT / "include" / "c3l_hip.h").read_text()
    body = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(c3l_[a-z0-9_]+)\s*\(", body)))


def test_header_declares_the_hot_path():
    names = declared_symbols()
    for required in ("c3l_plan_create", "c3l_plan_load", "c3l_plan_apply", "c3l_tilde_unpack", "c3l_zmultidot",
                     "c3l_shifted_triangular_green"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from c3lindblad_amd import _lib

    lib = _lib.load_library()
    missing = [name for name in declared_symbols() if not hasattr(lib, name)]
    assert not missing, missing
    assert set(_lib.SIGNATURES) >= set(declared_symbols())


def test_invalid_arguments_are_rejected_before_any_device_work():
    from c3lindblad_amd import _lib

    lib = _lib.load_library()
    handle = ctypes.c_void_p()
    status = lib.c3l_plan_create(0, 0, None, 0, 0, None, None, None, ctypes.byref(handle))
    assert status == _lib.C3L_EINVAL
    assert b"invalid" in lib.c3l_last_error()
    with pytest.raises(ValueError):
        _lib.check(status)
    assert lib.c3l_version().startswith(b"c3l_hip")

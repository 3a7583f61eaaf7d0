"""Compile ``csrc/c3l_hip.hip`` into ``libc3l_hip.so`` for gfx950 with hipcc."""

from __future__ import annotations

import os
import subprocess
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
SOURCES = sorted((PACKAGE / "csrc").glob("*.hip"))
TARGET = PACKAGE / "libc3l_hip.so"

HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "--offload-arch=gfx950",
    # packed FP32 instructions stay off for every kernel in this library
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]


def build(verbose: bool = False) -> Path:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = TARGET.with_suffix(".so.tmp")
    cmd = [hipcc, *HIPCC_FLAGS, "-o", str(tmp), *[str(s) for s in SOURCES]]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError(f"hipcc failed ({result.returncode}):\n{result.stderr}")
    if verbose and result.stderr:
        print(result.stderr)
    os.replace(tmp, TARGET)
    return TARGET


if __name__ == "__main__":
    print(build(verbose=True))

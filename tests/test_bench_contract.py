"""The benchmark's command-line contract on the host: a multi-GPU request ends with a message and
exit status 2 before any GPU work (the path runs on one MI355X)."""

import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.parametrize("argv, env", [(["--gpus", "2"], {}), (["--gpus", "1"], {"WORLD_SIZE": "2"})])
def test_bench_refuses_more_than_one_gpu(argv, env):
    run = subprocess.run([sys.executable, str(ROOT / "bench.py"), *argv], cwd=ROOT, capture_output=True, text=True,
                         timeout=300, env={**os.environ, **env})
    assert run.returncode == 2, run.stderr
    assert "single-GPU path only" in run.stderr
    assert run.stdout.strip() == ""  # no JSON line for a run that did not happen

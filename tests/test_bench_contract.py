"""bench.py's command-line contract (CPU): one GPU only -- --gpus above 1,
or a multi-rank launch, exits with a message before touching a device."""
from __future__ import annotations

import os
import subprocess
import sys

import oracle_lib as O


def run(args, env=None):
    return subprocess.run([sys.executable, str(O.ROOT / "bench.py")] + args, capture_output=True, text=True,
                          timeout=120, env=env)


def test_more_than_one_gpu_is_refused():
    out = run(["--gpus", "2", "--steps", "1", "--warmup", "0"])
    assert out.returncode == 2
    assert "one GPU" in out.stderr and out.stdout == ""
    env = dict(os.environ, WORLD_SIZE="4")
    out = run(["--steps", "1"], env=env)
    assert out.returncode == 2


def test_unknown_workload_is_an_argument_error():
    out = run(["--workload", "nope"])
    assert out.returncode == 2 and "invalid choice" in out.stderr

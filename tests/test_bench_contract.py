"""bench.py's contract on the CPU: one GPU only -- --gpus above 1, or a
multi-rank launch, exits with a message before touching a device -- and the
roofline figures of its JSON line (algorithmic bytes over the measured
duration, and the HBM traffic read from the committed PMC summaries)."""
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


def _bench_module():
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_mod", O.ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_roofline_arithmetic():
    # 12 nnz + 40 n bytes of the 1100 x 1100 FEM system (nnz 8,430,446,
    # n 1,205,604) in 22.199 us: 149,389,512 B / 22.199e-6 s = 6,729.6 GB/s,
    # 84.12 % of 8,000 GB/s.
    B = _bench_module()
    r = B.roof("k", 12 * 8_430_446 + 40 * 1_205_604, 22.199, "")
    assert r["alg_bytes"] == 149_389_512
    assert r["achieved"] == 6729.6 and r["frac"] == 0.8412
    assert r["bound"] == "hbm" and r["peak"] == 8000.0 and r["unit"] == "GB/s"
    # 1 GB in 250 us is 4,000 GB/s
    assert B.roof("k", 1e9, 250.0, "")["achieved"] == 4000.0


def test_traffic_comes_from_the_committed_pmc_passes():
    B = _bench_module()
    prof = B.profiled_traffic()
    assert prof["k_pcg_spmv_fused_acc"][0] == 134_989_056
    assert prof["k_pcg_update_xrz_acc"][0] == 61_294_143
    assert all(src.startswith("profiles/") for _, src in prof.values())


def test_break_even_right_hand_sides():
    """The learned method's set-up costs more and its solve less: it is ahead
    from the first k with S_l + k T_l <= S_b + k T_b."""
    B = _bench_module()
    assert B.break_even({"setup_ms": 14.0, "solve_ms": 240.0}, {"setup_ms": 5.0, "solve_ms": 520.0}) == 1
    assert B.break_even({"setup_ms": 900.0, "solve_ms": 240.0}, {"setup_ms": 5.0, "solve_ms": 520.0}) == 4
    assert B.break_even({"setup_ms": 14.0, "solve_ms": 600.0}, {"setup_ms": 5.0, "solve_ms": 520.0}) is None

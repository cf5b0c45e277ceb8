"""Run reports (CPU): the HTML view of a report JSON with the reference's
keys (archived/host/report_io.hpp:165-273), and xplus_device's argument
errors (archived/host/xrt_host.cpp:115-175: usage, exit code 1)."""
from __future__ import annotations

import json
import subprocess

import numpy as np

import xplus_amd as X
from xplus_amd import report as R


def sample(n_iter=300, ok=True):
    res = np.logspace(0, -6, n_iter)
    return {
        "dataset": {"path": "data/x<y>", "xclbin": "mi355x", "n": 512, "nnz": 2464, "max_iters": 2048,
                    "tau": 1e-10, "device_index": 0},
        "result": {"pass": ok, "converged": ok, "status": "converged" if ok else "max_iter",
                   "iterations": n_iter, "final_rr": float(res[-1] ** 2), "final_residual_norm": float(res[-1]),
                   "residual_rel": 1e-7, "solution_max_abs_diff": 1e-9, "golden_residual_norm": 1e-6},
        "host_timing_ms": {"total": 10.0, "host_setup": 2.0, "buffer_h2d": 0.5, "kernel_total": 6.0,
                           "buffer_d2h": 0.5, "verify": 1.0},
        "kernel_timing_ms": {"pcg_control_total": 6.0, "pcg_control_calls": 1, "spmv_total": 3.0,
                             "spmv_calls": n_iter, "update_xrz_total": 2.0, "update_xrz_calls": n_iter,
                             "split_timing_available": True, "timing_model": "split_kernels"},
        "trace_metadata": {"source": "hardware", "note": "device trace"},
        "iterations_trace": [{"iteration": k, "alpha": 0.5, "beta": 0.1, "rz": float(r), "rr": float(r * r),
                              "residual": float(r)} for k, r in enumerate(res)],
    }


def test_html_render(tmp_path):
    js = tmp_path / "r.json"
    js.write_text(json.dumps(sample()))
    assert R.main(["static", str(js), str(tmp_path / "r.html")]) == 0
    page = (tmp_path / "r.html").read_text()
    assert "PASS" in page and "<polyline" in page and "data/x&lt;y&gt;" in page
    assert page.count("<tr><td>") >= 100                       # first and last 50 iterations
    assert "... 200 iterations ..." in page
    js.write_text(json.dumps(sample(n_iter=1, ok=False)))
    assert R.main([str(js), str(tmp_path / "f.html")]) == 0
    page = (tmp_path / "f.html").read_text()
    assert "FAIL" in page and "No per-iteration trace" in page
    assert R.main(["only-one-arg"]) == 2


def test_device_cli_rejects_bad_arguments(tmp_path):
    for args in (["--tau", "x"], ["--tau", "-1"], ["--max-iters", "-2"], ["--bogus"], ["--json-out"]):
        out = subprocess.run([str(X.DEVICE_CLI_PATH)] + args, capture_output=True, text=True, timeout=60)
        assert out.returncode == 1 and "Usage:" in out.stderr, (args, out.stderr)

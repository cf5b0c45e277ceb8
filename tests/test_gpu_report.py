"""xplus_device (the whole-solve device path behind the reference FPGA
host's command line) and its reports, on the reference's n512 dataset.

Bars: the text and JSON reports have the reference's sections and keys;
pass / exit code agree with the reference's rule (converged and
max |x - x_golden| <= 1e-8 -> 0, else 3; not converged -> 2); the
per-iteration trace has one row per iteration (0-based), its last rr is the
report's final_rr, and its rr follows the oracle's FP64 golden trace within
relative 1e-3 while rr > 1e-8 rr_0; the HTML view renders from the JSON.
"""
from __future__ import annotations

import json
import subprocess

import numpy as np
import pytest

import oracle_lib as O
import xplus_amd as X
from xplus_amd import report as R

pytestmark = pytest.mark.gpu
N512 = O.ROOT / "tests" / "golden" / "n512"


def test_device_cli_reports(gpu, tmp_path):
    js, txt = tmp_path / "rep" / "r.json", tmp_path / "rep" / "r.txt"
    out = subprocess.run([str(X.DEVICE_CLI_PATH), "mi355x", str(N512), "--tau", "1e-10", "--json-out", str(js),
                          "--txt-out", str(txt)], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode in (0, 3), out.stderr
    data = json.loads(js.read_text())
    res = data["result"]
    assert res["converged"] and res["status"] == "converged"
    assert res["pass"] == (res["solution_max_abs_diff"] <= 1e-8)
    assert out.returncode == (0 if res["pass"] else 3)
    assert "[done] iter=" in out.stdout and "[host-timing-ms]" in out.stdout and "[kernel-timing-ms]" in out.stdout
    tr = data["iterations_trace"]
    assert len(tr) == res["iterations"] and [t["iteration"] for t in tr] == list(range(len(tr)))
    assert tr[-1]["rr"] == pytest.approx(res["final_rr"], rel=1e-11)
    a = X.load_dataset(N512)
    g = O.golden(a, tau=1e-10)
    rr_g = g["trace"][:, 4]
    rr0 = np.linalg.norm(a.b - O.spmv(a, a.x0)) ** 2
    k = min(len(tr), len(rr_g))
    rr = np.array([t["rr"] for t in tr[:k]])
    early = rr_g[:k] > 1e-8 * rr0
    np.testing.assert_allclose(rr[early], rr_g[:k][early], rtol=1e-3)
    assert data["kernel_timing_ms"]["split_timing_available"] is True
    assert data["trace_metadata"]["source"] == "hardware"
    text = txt.read_text()
    for section in ("[Problem]", "[Accuracy]", "[Host Timing ms]", "[Kernel Timing ms]"):
        assert section in text
    assert f"iterations         : {res['iterations']}" in text
    R.render_file(js, tmp_path / "r.html")
    assert "<polyline" in (tmp_path / "r.html").read_text()

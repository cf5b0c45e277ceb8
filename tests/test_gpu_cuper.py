"""GPU runs of matrices in the reference's Cuper packed format and of the
Cuper command line (xplus_cuper).

Bars (written in each test):
  * a matrix packed by the reference's own packer (oracle/_ref) -- or by the
    library's, which equals it word for word (tests/test_cuper_format.py) --
    imported and built into a device matrix: y = A b in FP32 within
    (k+1) u sum|a32 b32| per row of the FP32-valued product (u = 2^-24), and
    within the reference's diff_tol 1e-3 of the FP64 product (its
    --spmv-only check, host/cuper_tapa_pcg_main.cpp:349-386);
  * xplus_cuper --spmv-only prints the reference's [done] / [check] /
    [timing-ms] lines, exit code 0, differences within diff_tol;
  * xplus_cuper (PCG mode) on the config-1 dataset converges and its
    solution is within diff_tol of the FP64 multi-kernel solver's.
"""
from __future__ import annotations

import re
import subprocess

import numpy as np
import pytest

import oracle_lib as O
import xplus_amd as X

pytestmark = pytest.mark.gpu
U32 = 2.0 ** -24


def packed_of(a, tmp_path):
    if O.ref is not None and hasattr(O.ref, "ref_cuper_pack_info"):
        a.save(tmp_path)
        ptr, data = O.ref_cuper_pack(tmp_path)
        return X.CuperPacked(ptr, data), "reference packer"
    return X.cuper_pack(a), "library packer"


@pytest.mark.parametrize("which", ["c1", "fem_multi_batch"])
def test_spmv_on_imported_packing(gpu, tmp_path, which):
    a = X.gen_diffusion2d(32, 32, 1.0, 20261204) if which == "c1" else X.gen_fem2d(150, 1.0, 4, 1e2, 0.3, 8)
    p, src = packed_of(a, tmp_path)
    imp = p.unpack(a.n)
    print(f"{which}: n={a.n} packed by the {src}: batches={p.batch_count} matrix_len={p.matrix_len}", flush=True)
    m = X.Matrix(imp)
    info = m.info()
    perm = m.permutation()
    real = perm >= 0
    b = a.b
    bn = X.DeviceArray.from_numpy(b)
    xs = X.DeviceArray(info["n_pad"], np.float32)
    ys = X.DeviceArray(info["n_pad"], np.float32)
    m.to_solve_order(bn.ptr, xs.ptr)
    m.spmv_f32(xs.ptr, ys.ptr)
    X.synchronize()
    y = np.zeros(a.n)
    y[perm[real]] = ys.numpy()[real]
    b32 = b.astype(np.float32).astype(np.float64)
    ref32 = O.spmv(imp, b32)
    absrow = O.spmv(X.Csr(imp.row_ptr, imp.col_idx, np.abs(imp.values)), np.abs(b32))
    k = np.diff(imp.row_ptr).max() + 1
    assert np.all(np.abs(y - ref32) <= k * U32 * absrow)
    exact = O.spmv(a, b)
    d_abs = np.abs(y - exact)
    d_rel = d_abs / np.maximum(np.abs(exact), 1e-12)
    print(f"{which}: max_abs_diff {d_abs.max():.3e} max_rel_diff {d_rel.max():.3e}", flush=True)
    assert not (d_abs.max() > 1e-3 and d_rel.max() > 1e-3)


def run_cli(args):
    out = subprocess.run([str(X.CUPER_CLI_PATH)] + [str(v) for v in args], capture_output=True, text=True,
                         timeout=120)
    print(out.stdout, out.stderr, flush=True)
    return out


def fields(line):
    return {k: v for k, v in re.findall(r"(\w+)=(\S+)", line)}


def test_cli_spmv_only(gpu, tmp_path):
    a = X.gen_fem2d(150, 1.0, 4, 1e2, 0.3, 8)
    a.save(tmp_path)
    out = run_cli([tmp_path, "--spmv-only", "--spmv-repeats", 20])
    assert out.returncode == 0
    lines = {ln.split("]")[0] + "]": ln for ln in out.stdout.splitlines()}
    assert "mode=cuper-spmv-tapa" in lines["[xplus]"]
    done = fields(lines["[done]"])
    assert done["mode"] == "spmv_only" and done["spmv_calls"] == "20" and done["status"] == "ok"
    chk = fields(lines["[check]"])
    assert float(chk["max_abs_diff"]) <= 1e-3 or float(chk["max_rel_diff"]) <= 1e-3
    tm = fields(lines["[timing-ms]"])
    assert float(tm["spmv_total"]) > 0 and float(tm["gflops"]) > 0
    assert abs(float(tm["spmv_avg"]) * 20 - float(tm["spmv_total"])) <= 1e-9 * float(tm["spmv_total"]) + 1e-12
    # a tolerance nothing meets: exit code 3, as the reference
    out = run_cli([tmp_path, "--spmv-only", "--diff-tol", "1e-300"])
    assert out.returncode == 3


def test_cli_pcg_mode(gpu, tmp_path):
    a = X.gen_diffusion2d(32, 32, 1.0, 20261204)
    a.save(tmp_path)
    out = run_cli([tmp_path, "--tau", "1e-10", "--diff-tol", "1e-3"])
    assert out.returncode == 0
    lines = {ln.split("]")[0] + "]": ln for ln in out.stdout.splitlines()}
    done = fields(lines["[done]"])
    assert done["status"] == "converged" and int(done["iter"]) > 0
    chk = fields(lines["[check]"])
    assert float(chk["max_abs_diff"]) <= 1e-3 or float(chk["max_rel_diff"]) <= 1e-3
    assert float(fields(lines["[timing-ms]"])["spmv_avg"]) > 0

"""Pins the oracle (oracle/xplus_oracle.c) before it is trusted as the checker.

Each oracle function is checked against at least one of:
  * the reference's own code compiled from its sources (oracle/_ref, built by
    oracle/Makefile in the build container) -- bitwise;
  * the reference's logged known answers for the generated n512 dataset
    (docs/bitstream_summaries/2026-07-07-cuper-tapa-pcg-callipepla/testing.md:15-37,
    docs/design/implementation_versions_zh.md:317-330);
  * closed forms / textbook routines (dense matvec, Cholesky solve, diagonal
    systems, CG finite termination);
chosen so that a dropped term, a wrong sign or index, or a transposed
operand fails.
"""
from __future__ import annotations

import ctypes as C
import os
import shutil
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O
import xplus_amd as X

GOLDEN = Path(__file__).resolve().parent / "golden"
N512 = GOLDEN / "n512"
needs_ref = pytest.mark.skipif(O.ref is None, reason="reference build oracle/_ref not present")


def small_spd(n=12, seed=0, density=0.3):
    rng = np.random.default_rng(seed)
    m = np.where(rng.random((n, n)) < density, rng.integers(-3, 4, (n, n)), 0).astype(float)
    a = -np.abs(m + m.T)
    np.fill_diagonal(a, 0.0)
    np.fill_diagonal(a, -a.sum(1) + 1.0 + rng.integers(0, 3, n))
    rp, ci, va = [0], [], []
    for r in range(n):
        for c in range(n):
            if a[r, c] != 0.0:
                ci.append(c)
                va.append(a[r, c])
        rp.append(len(ci))
    return X.Csr(rp, ci, va, rng.integers(-5, 6, n).astype(float), np.zeros(n)), a


# ------------------------------------------------------------------ loader
def write_ds(d: Path, rp, ci, va, b, x0):
    d.mkdir(parents=True, exist_ok=True)
    for name, arr in (("row_ptr", rp), ("col_idx", ci), ("values", va), ("b", b), ("x0", x0)):
        (d / f"{name}.txt").write_text(" ".join(str(v) for v in arr) + "\n")


LOADER_CASES = [
    ("row_ptr.txt must contain at least 2 entries", dict(rp=[0])),
    ("col_idx.txt and values.txt size mismatch", dict(va=[1.0])),
    ("row_ptr[0] must be 0", dict(rp=[1, 2, 2])),
    ("row_ptr[n] must equal nnz", dict(rp=[0, 1, 3])),
    ("row_ptr must be nondecreasing", dict(rp=[0, 2, 1, 2], b=[1, 1, 1], x0=[0, 0, 0])),
    ("col_idx out of range", dict(ci=[0, 2])),
    ("b.txt length must equal n", dict(b=[1.0])),
    ("x0.txt length must equal n", dict(x0=[0.0, 0.0, 0.0])),
]


@pytest.mark.parametrize("msg,override", LOADER_CASES)
def test_loader_messages(tmp_path, msg, override):
    base = dict(rp=[0, 1, 2], ci=[0, 1], va=[2.0, 3.0], b=[1.0, 1.0], x0=[0.0, 0.0])
    base.update(override)
    write_ds(tmp_path, base["rp"], base["ci"], base["va"], base["b"], base["x0"])
    with pytest.raises(O.OracleError, match=msg.replace("[", r"\[").replace("]", r"\]")):
        O.load(tmp_path)
    # the product loader raises the same message
    with pytest.raises(X.XplusError) as e:
        X.load_dataset(tmp_path)
    assert str(e.value) == msg


def test_loader_missing_and_empty(tmp_path):
    with pytest.raises(O.OracleError, match="failed to open"):
        O.load(tmp_path / "nope")
    with pytest.raises(X.XplusError, match="failed to open"):
        X.load_dataset(tmp_path / "nope")
    write_ds(tmp_path, [0, 1], [0], [2.0], [1.0], [0.0])
    (tmp_path / "values.txt").write_text("  \n")
    with pytest.raises(O.OracleError, match="empty array file"):
        O.load(tmp_path)
    with pytest.raises(X.XplusError, match="empty array file"):
        X.load_dataset(tmp_path)


def test_loader_stops_at_first_bad_token(tmp_path):
    # `while (input >> value)` stops silently at the first unparsable token.
    write_ds(tmp_path, [0, 1, 2], [0, 1], [2.0, 3.0], [1.0, 1.0], [0.0, 0.0])
    (tmp_path / "b.txt").write_text("1.5 2.5 junk 7\n")
    d = O.load(tmp_path)
    np.testing.assert_array_equal(d["b"], [1.5, 2.5])
    np.testing.assert_array_equal(X.load_dataset(tmp_path).b, [1.5, 2.5])


@needs_ref
def test_loader_matches_reference_n512():
    d = O.load(N512)
    n, nnz = C.c_int(), C.c_int()
    err = C.create_string_buffer(256)
    assert O.ref.ref_dataset_info(str(N512).encode(), C.byref(n), C.byref(nnz), err, 256) == 0
    rp = np.zeros(n.value + 1, np.int32)
    ci = np.zeros(nnz.value, np.int32)
    va, b, x0 = np.zeros(nnz.value), np.zeros(n.value), np.zeros(n.value)
    assert O.ref.ref_dataset_read(str(N512).encode(), rp, ci, va, b, x0, err, 256) == 0
    for k, v in (("row_ptr", rp), ("col_idx", ci), ("values", va), ("b", b), ("x0", x0)):
        np.testing.assert_array_equal(d[k], v)
    p = X.load_dataset(N512)
    np.testing.assert_array_equal(p.values, va)
    np.testing.assert_array_equal(p.b, b)


# ------------------------------------------------------------------ SpMV / diag
def test_spmv_integer_exact():
    a, dense = small_spd(20, 3)
    x = np.arange(20, dtype=float) - 7.0
    np.testing.assert_array_equal(O.spmv(a, x), dense @ x)


def test_extract_diag_errors():
    a = X.Csr([0, 1, 2], [1, 0], [1.0, 1.0])
    err = C.create_string_buffer(128)
    d = np.zeros(2)
    assert O.oracle.orc_extract_jacobi_diag(2, a.row_ptr, a.col_idx, a.values, d, err, 128) != 0
    assert err.value.decode() == "missing diagonal entry at row 0"
    a = X.Csr([0, 1, 2], [0, 1], [2.0, 0.0])
    assert O.oracle.orc_extract_jacobi_diag(2, a.row_ptr, a.col_idx, a.values, d, err, 128) != 0
    assert err.value.decode() == "zero diagonal entry at row 1"
    # the first diagonal entry wins when duplicated
    a = X.Csr([0, 2], [0, 0], [3.0, 4.0])
    assert O.oracle.orc_extract_jacobi_diag(1, a.row_ptr, a.col_idx, a.values, d, err, 128) == 0
    assert d[0] == 3.0


# ------------------------------------------------------------------ golden PCG
def test_golden_known_answers_n512():
    """Reference logs for data/generated/cgsolver/n512 (FP32-SpMV path, so ~1e-6)."""
    d = O.load(N512)
    a = X.Csr(d["row_ptr"], d["col_idx"], d["values"], d["b"], d["x0"])
    g0 = O.golden(a, tau=1e300)
    assert g0["iterations"] == 0
    assert g0["final_rr"] == pytest.approx(3.427876635047e+01, rel=1e-6)
    g1 = O.golden(a, tau=1e-30, max_iters=1)
    assert g1["trace"][0, 4] == pytest.approx(9.407299557997e+01, rel=1e-6)
    assert O.residual_norm(a, g1["x"]) == pytest.approx(9.699123342347, rel=1e-6)
    g = O.golden(a, tau=1e-8, max_iters=100)
    assert g["iterations"] == 41 and g["converged"]


def test_golden_diagonal_one_iteration():
    n = 7
    diag = np.arange(1, n + 1, dtype=float)
    a = X.Csr(np.arange(n + 1), np.arange(n), diag, np.linspace(-1, 1, n), np.zeros(n))
    g = O.golden(a, tau=1e-24)
    assert g["iterations"] == 1
    np.testing.assert_allclose(g["x"], a.b / diag, rtol=1e-15)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_golden_matches_cholesky_and_terminates(seed):
    a, dense = small_spd(16, seed)
    g = O.golden(a, tau=1e-26, max_iters=200)
    assert g["converged"] and g["iterations"] <= 16 + 2
    L = np.linalg.cholesky(dense)
    xs = np.linalg.solve(L.T, np.linalg.solve(L, a.b))
    np.testing.assert_allclose(g["x"], xs, rtol=1e-10, atol=1e-12)
    tr = g["trace"]
    assert np.all(tr[:, 3] > 0)                       # rz > 0 on SPD
    np.testing.assert_allclose(tr[:, 5], np.sqrt(tr[:, 4]))
    np.testing.assert_array_equal(tr[:, 0], np.arange(len(tr)))


def test_golden_errors_and_conjugacy():
    a, dense = small_spd(10, 5)
    with pytest.raises(O.OracleError, match="tau must be positive"):
        O.golden(a, tau=0.0)
    # beta recorded on the converging iteration too, break before p update
    g = O.golden(a, tau=1e-20, max_iters=3)
    assert g["trace"].shape[0] == 3 and np.all(np.isfinite(g["trace"][:, 2]))
    # max_iters = 0 means max(4n, 1000)
    g = O.golden(a, tau=1e-300, max_iters=0, trace_cap=2000)
    assert g["iterations"] <= 1000


def test_golden_first_step_closed_form():
    """One iteration written out by hand: catches sign/index slips in the loop."""
    a, dense = small_spd(9, 7)
    b = a.b
    d = np.diag(dense)
    r = b.copy()
    z = r / d
    p = z.copy()
    rz = r @ z
    ap = dense @ p
    alpha = rz / (p @ ap)
    x1 = alpha * p
    r1 = r - alpha * ap
    z1 = r1 / d
    g = O.golden(a, tau=1e-300, max_iters=1)
    np.testing.assert_allclose(g["x"], x1, rtol=1e-13)
    np.testing.assert_allclose(g["trace"][0, 1:5], [alpha, (r1 @ z1) / rz, r1 @ z1, r1 @ r1], rtol=1e-13)


@needs_ref
@pytest.mark.parametrize("case", ["n512", "c1"])
def test_golden_bitwise_vs_reference(case, tmp_path):
    if case == "n512":
        ddir = N512
        d = O.load(ddir)
        a = X.Csr(d["row_ptr"], d["col_idx"], d["values"], d["b"], d["x0"])
        tau = 1e-10
    else:
        a = X.gen_diffusion2d(32, 32, 1.0, 20261204)
        ddir = tmp_path / "c1"
        a.save(ddir)
        a = X.load_dataset(ddir)  # the %.17g text round trip is exact
        tau = 1e-12 * float(a.b @ a.b)
    g = O.golden(a, tau=tau)
    x = np.zeros(a.n)
    tr = np.zeros(6 * 5000)
    it, cv = C.c_int(), C.c_int()
    sc = np.zeros(3)
    err = C.create_string_buffer(256)
    assert O.ref.ref_golden(str(ddir).encode(), tau, 0, x, tr, 5000, C.byref(it), C.byref(cv), sc, err, 256) == 0
    assert it.value == g["iterations"]
    np.testing.assert_array_equal(x, g["x"])
    np.testing.assert_array_equal(tr[: 6 * it.value].reshape(-1, 6), g["trace"])
    assert sc[1] == O.residual_norm(a, g["x"])


# ------------------------------------------------------------------ six kernels + multi-kernel
@needs_ref
def test_six_kernels_bitwise_vs_reference():
    rng = np.random.default_rng(11)
    n = 700
    a = X.gen_diffusion2d(28, 25, 1.0, 3)
    x = rng.standard_normal(a.n)
    y_o, y_r = np.zeros(a.n), np.zeros(a.n)
    O.oracle.orc_k_spmv_csr(a.row_ptr, a.col_idx, a.values, x, y_o, a.n)
    O.ref.spmv_csr_kernel(a.row_ptr, a.col_idx, a.values, x, y_r, a.n)
    np.testing.assert_array_equal(y_o, y_r)
    vecs = [rng.standard_normal(n) for _ in range(6)]
    outs_o = [v.copy() for v in vecs]
    outs_r = [v.copy() for v in vecs]
    mo, mr = np.zeros(2), np.zeros(2)
    O.oracle.orc_k_init_pcg(vecs[0], vecs[1], vecs[2], outs_o[3], outs_o[4], outs_o[5], mo, n)
    O.ref.init_pcg_kernel(vecs[0], vecs[1], vecs[2], outs_r[3], outs_r[4], outs_r[5], mr, n)
    for k in (3, 4, 5):
        np.testing.assert_array_equal(outs_o[k], outs_r[k])
    np.testing.assert_array_equal(mo, mr)
    O.oracle.orc_k_dot(vecs[0], vecs[1], mo, n)
    O.ref.dot_kernel(vecs[0], vecs[1], mr, n)
    assert mo[0] == mr[0]
    outs_o = [v.copy() for v in vecs]
    outs_r = [v.copy() for v in vecs]
    O.oracle.orc_k_update_xrz(outs_o[0], outs_o[1], outs_o[2], outs_o[3], outs_o[4], outs_o[5], mo, 0.37, n)
    O.ref.update_xrz_kernel(outs_r[0], outs_r[1], outs_r[2], outs_r[3], outs_r[4], outs_r[5], mr, 0.37, n)
    for k in range(6):
        np.testing.assert_array_equal(outs_o[k], outs_r[k])
    np.testing.assert_array_equal(mo, mr)
    O.oracle.orc_k_update_p(outs_o[0], outs_o[1], -1.25, n)
    O.ref.update_p_kernel(outs_r[0], outs_r[1], -1.25, n)
    np.testing.assert_array_equal(outs_o[1], outs_r[1])


def test_six_kernels_closed_forms():
    rng = np.random.default_rng(12)
    n = 3000  # beyond the reference's kMaxN: the oracle has no cap
    b, ax, mi, x, p, r, ap = (rng.standard_normal(n) for _ in range(7))
    r_, z_, p_ = np.zeros(n), np.zeros(n), np.zeros(n)
    m = np.zeros(2)
    O.oracle.orc_k_init_pcg(b, ax, mi, r_, z_, p_, m, n)
    np.testing.assert_allclose(r_, b - ax, rtol=0)
    np.testing.assert_allclose(z_, mi * (b - ax), rtol=0)
    np.testing.assert_array_equal(p_, z_)
    np.testing.assert_allclose(m, [r_ @ z_, r_ @ r_], rtol=1e-12)
    x0, r0 = x.copy(), r.copy()
    z = np.zeros(n)
    O.oracle.orc_k_update_xrz(x, p, r, ap, mi, z, m, 0.5, n)
    np.testing.assert_allclose(x, x0 + 0.5 * p, rtol=1e-15)
    np.testing.assert_allclose(r, r0 - 0.5 * ap, rtol=1e-15)
    np.testing.assert_allclose(z, mi * r, rtol=1e-15)
    # negative n is a silent no-op
    before = r.copy()
    O.oracle.orc_k_update_xrz(x, p, r, ap, mi, z, m, 0.5, -1)
    np.testing.assert_array_equal(r, before)


@needs_ref
def test_multi_kernel_vs_reference_n512():
    d = O.load(N512)
    a = X.Csr(d["row_ptr"], d["col_idx"], d["values"], d["b"], d["x0"])
    mk = O.multi_kernel(a, tau=1e-8, max_iters=100)
    x = np.zeros(a.n)
    it, st, emi = C.c_int(), C.c_int(), C.c_int()
    sc = np.zeros(3)
    err = C.create_string_buffer(256)
    assert O.ref.ref_multi_kernel(str(N512).encode(), 1e-8, 100, x, C.byref(it), C.byref(st), C.byref(emi), sc,
                                  err, 256) == 0
    assert (it.value, st.value, emi.value) == (mk["iterations"], mk["status"], mk["effective_max_iters"]) == (41, 0, 100)
    np.testing.assert_array_equal(x, mk["x"])
    np.testing.assert_array_equal(sc, [mk["final_rr"], mk["residual_l2"], mk["residual_rel"]])


def test_multi_kernel_nan_rhs_is_max_iter():
    """rr = NaN fails `rr > tau`, so the reference loop never runs: status max_iter
    (archived/host/multi_kernel_solver.hpp:156-158, 227-231)."""
    a, _ = small_spd(8, 1)
    b = a.b.copy()
    b[3] = np.nan
    mk = O.multi_kernel(a, tau=1e-10, b=b)
    assert mk["status"] == 1 and mk["iterations"] == 0


@needs_ref
def test_cuper_style_spmv_vs_reference(tmp_path):
    a = X.gen_fem2d(130, 1.0, 5, 1e3, 0.3, 9)  # n = 16384 > 8192: two column batches
    a.save(tmp_path)
    a = X.load_dataset(tmp_path)
    x = np.random.default_rng(4).standard_normal(a.n)
    y_o, y_r = np.zeros(a.n), np.zeros(a.n)
    O.oracle.orc_cuper_style_spmv(a.n, a.row_ptr, a.col_idx, a.values, x, y_o)
    err = C.create_string_buffer(256)
    assert O.ref.ref_cuper_style_spmv(str(tmp_path).encode(), x, y_r, err, 256) == 0
    np.testing.assert_array_equal(y_o, y_r)
    # and it is the FP32-product SpMV within the rounding bound
    y64 = O.spmv(a, x)
    bound = 1e-6 * O.spmv(X.Csr(a.row_ptr, a.col_idx, np.abs(a.values)), np.abs(x)) + 1e-300
    assert np.all(np.abs(y_o - y64) <= bound)


@needs_ref
def test_reference_executable_on_n512():
    """The reference's own local executable passes on our restated n512 dataset."""
    import subprocess
    out = subprocess.run([str(O.REF_BIN), str(N512), "--tau", "1e-8", "--max-iters", "100"],
                         capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "[done] iter=41" in out.stdout

"""CPU-side tests of the product library: the C-ABI exports, the host-side
SELL-C-sigma builder, the generators and the loader.  No GPU compute here."""
from __future__ import annotations

import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O
import xplus_amd as X

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "xplus_capi.h"


def declared_functions() -> list[str]:
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", text, flags=re.M)
    return sorted({n for n in names if n not in ("if", "extern")})


def test_header_declares_the_reference_surface():
    names = declared_functions()
    for ref_name in ("spmv_csr_kernel", "spmv_blocked_kernel", "init_pcg_kernel", "dot_kernel",
                     "update_xrz_kernel", "update_p_kernel"):
        assert ref_name in names
    assert len(names) > 30


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", str(X.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing


def test_no_gpu_fails_loudly():
    if X.device_count() > 0:
        pytest.skip("GPU present")
    a = X.gen_diffusion2d(8, 8, 1.0, 1)
    with pytest.raises(X.XplusError, match="no GPU visible"):
        X.Matrix(a)
    # the void drop-in functions record the error instead of computing on the CPU
    y = np.full(a.n, 7.0)
    X.spmv_csr_kernel(a.row_ptr, a.col_idx, a.values, np.ones(a.n), y, a.n)
    assert np.all(y == 7.0)
    assert "no GPU visible" in X.last_error()


def test_device_ops_reject_null_arguments():
    """The stream-ordered FP32 operations check their pointers before any
    device work (no GPU needed): a null matrix or vector is an error."""
    import xplus_amd as XA
    assert XA._op_dot(None, None, None, None, None) == -1
    assert "null argument" in X.last_error()
    assert XA._op_init(None, None, None, None, None, None, None, None, None) == -1
    assert XA._op_xrz(None, None, None, None, None, None, None, None, 0.5, None) == -1
    assert XA._op_p(None, None, None, 0.5, None) == -1
    assert "null argument" in X.last_error()


# ------------------------------------------------------------------ generators
def test_diffusion2d_structure():
    a = X.gen_diffusion2d(32, 32, 1.0, 20261204)
    assert (a.n, a.nnz) == (1024, 4992)           # 5 N^2 - 4 N
    A = a.to_scipy()
    assert abs(A - A.T).max() == 0.0
    d = A.diagonal()
    off = np.asarray(abs(A).sum(axis=1)).ravel() - d
    assert np.all(d >= off * (1 - 1e-14))         # weakly dominant: interior rows sum to zero
    assert np.sum(d > off * (1 + 1e-12)) == 4 * 32 - 4   # boundary rows: ghost edges add to the diagonal
    assert np.all(np.diff(a.col_idx[a.row_ptr[0]:a.row_ptr[1]]) > 0)
    assert a.b.min() >= -1 and a.b.max() <= 1 and np.all(a.x0 == 0)
    b2 = X.gen_diffusion2d(32, 32, 1.0, 20261204)
    np.testing.assert_array_equal(a.values, b2.values)   # deterministic
    c = X.gen_diffusion2d(32, 32, 1.0, 20261205)
    assert not np.array_equal(a.values, c.values)


def test_diffusion3d_structure():
    k = 10
    a = X.gen_diffusion3d(k, 1.5, 3)
    assert (a.n, a.nnz) == (k ** 3, 7 * k ** 3 - 6 * k ** 2)
    A = a.to_scipy()
    assert abs(A - A.T).max() == 0.0
    np.linalg.cholesky(A.toarray())


def test_fem2d_structure_and_spd():
    a = X.gen_fem2d(24, 1.0, 4, 1e3, 0.3, 7)
    assert a.n == 22 * 22
    A = a.to_scipy()
    assert abs(A - A.T).max() <= 1e-12 * abs(A).max()
    np.linalg.cholesky(A.toarray())              # SPD after Dirichlet elimination
    rl = np.diff(a.row_ptr)
    assert rl.min() >= 3 and rl.max() <= 9
    for r in range(a.n):
        cols = a.col_idx[a.row_ptr[r]:a.row_ptr[r + 1]]
        assert np.all(np.diff(cols) > 0)
    # thermal2-like density at scale: ~7 nonzeros per row
    big = X.gen_fem2d(202, 1.0, 20, 1e3, 0.3, 1)
    assert 6.6 < big.nnz / big.n < 7.2


def test_dataset_save_load_roundtrip(tmp_path):
    a = X.gen_fem2d(20, 1.0, 3, 1e3, 0.3, 2)
    a.b = np.random.default_rng(0).standard_normal(a.n)
    a.save(tmp_path)
    b = X.load_dataset(tmp_path)
    for k in ("row_ptr", "col_idx", "values", "b", "x0"):
        np.testing.assert_array_equal(getattr(a, k), getattr(b, k))
    o = O.load(tmp_path)
    np.testing.assert_array_equal(o["values"], a.values)


# ------------------------------------------------------------------ SELL builder
def sell_to_coo(s: dict, n: int):
    """Expands the SELL arrays back to (natural row, natural col, value) triples."""
    perm = s["perm"]
    rows, cols, vals = [], [], []
    sp = s["slice_ptr"]
    for sl in range(s["nslices"]):
        w = sp[sl + 1] - sp[sl]
        for j in range(w):
            e0 = (sp[sl] + j) * 64
            for lane in range(64):
                i = sl * 64 + lane
                if perm[i] < 0:
                    continue
                e = e0 + lane
                v = s["val"][e]
                if j == 0:
                    c = i
                elif s["idx16"]:
                    c = s["slice_base"][sl] + int(s["col16"][e])
                else:
                    c = int(s["col32"][e])
                if v != 0.0:
                    rows.append(perm[i])
                    cols.append(perm[c])
                    vals.append(v)
    return np.array(rows), np.array(cols), np.array(vals, np.float32)


@pytest.mark.parametrize("sigma", [64, 256, 1024])
def test_sell_roundtrip(sigma):
    a = X.gen_fem2d(40, 1.0, 5, 1e3, 0.3, 3)
    s = X.sell_build_host(a, sigma)
    assert s["n_pad"] % 64 == 0 and a.n <= s["n_pad"] < a.n + 64
    perm = s["perm"]
    real = perm[perm >= 0]
    assert sorted(real.tolist()) == list(range(a.n))       # a permutation
    # sorting stays inside sigma windows
    win = np.arange(len(perm)) // sigma
    assert np.all(real // sigma == win[perm >= 0]) or sigma >= a.n
    r, c, v = sell_to_coo(s, a.n)
    A = a.to_scipy().tocoo()
    got = {(int(i), int(j)): float(x) for i, j, x in zip(r, c, v)}
    want = {(int(i), int(j)): float(np.float32(x)) for i, j, x in zip(A.row, A.col, A.data) if x != 0.0}
    assert got == want
    assert s["idx16"] == 1


def test_sell_padding_is_small_with_sorting():
    a = X.gen_fem2d(202, 1.0, 20, 1e3, 0.3, 1)
    unsorted = X.sell_build_host(a, 64)
    sorted_ = X.sell_build_host(a, 1024)
    assert sorted_["stored"] < unsorted["stored"]
    assert sorted_["stored"] <= 1.03 * (a.nnz + (sorted_["n_pad"] - a.n))


def test_sell_handles_missing_diagonal_and_duplicates():
    # row 0: no diagonal; row 1: diagonal twice (both summed by SpMV)
    a = X.Csr([0, 1, 3, 4], [1, 1, 1, 2], [5.0, 2.0, 3.0, 1.0])
    s = X.sell_build_host(a, 64)
    r, c, v = sell_to_coo(s, a.n)
    trip = sorted(zip(r.tolist(), c.tolist(), v.tolist()))
    assert trip == [(0, 1, 5.0), (1, 1, 2.0), (1, 1, 3.0), (2, 2, 1.0)]


def test_sell_wide_span_uses_32bit_columns():
    n = 70000
    rp = np.arange(n + 1) * 2
    ci = np.empty(2 * n, np.int32)
    ci[0::2] = np.arange(n)
    ci[1::2] = (np.arange(n) + 69000) % n       # columns far from the row
    order = np.argsort(ci.reshape(n, 2), axis=1)
    ci = np.take_along_axis(ci.reshape(n, 2), order, axis=1).ravel()
    va = np.where(ci == np.repeat(np.arange(n), 2), 4.0, -1.0)
    a = X.Csr(rp, ci, va)
    s = X.sell_build_host(a, 64)
    assert s["idx16"] == 0


def test_sell_rejects_bad_csr():
    with pytest.raises(X.XplusError, match="col_idx out of range"):
        X.sell_build_host(X.Csr([0, 1], [3], [1.0]), 64)
    with pytest.raises(X.XplusError, match="sigma must be a multiple of 64"):
        X.sell_build_host(X.gen_diffusion2d(4, 4), 100)

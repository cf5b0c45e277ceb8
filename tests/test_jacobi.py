"""Plain Jacobi iteration: the FP32 oracle (orc_jacobi_prepare_f32 /
orc_jacobi_run_f32, the reference's PrepareJacobiVectorsAndR + RunJacobiCpu
restated) pinned by closed forms, and the library's host-side A * ones
right-hand side against it.  The reference's own host needs TAPA/XRT and
does not build here, so the pins are mathematical:

  * a diagonal system: one update gives fl(b * fl(1/d)), and then stays;
  * the error recursion: x_k = x* + (-D^-1 R)^k (x0 - x*) in FP64, for
    k = 1..8 and a random start (an off-by-one in the update count, a sign
    error or a diagonal left in R fails it);
  * the fixed point: many updates on a strictly diagonally dominant matrix
    reach numpy.linalg.solve within FP32 accuracy;
  * diff_max = max |x_k - x_{k-1}| by brute force from consecutive runs;
  * duplicate diagonal entries are summed into D;
  * A * ones on integer-valued matrices equals the exact row sums;
  * a tiny case in numpy float32 scalar arithmetic, bit for bit.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle_lib as O
import xplus_amd as X


def csr_of(A, dup_diag=None):
    import scipy.sparse as sp
    S = sp.csr_matrix(A)
    S.sort_indices()
    rp, ci, va = S.indptr.astype(np.int32), S.indices.astype(np.int32), S.data.astype(np.float64)
    if dup_diag:
        # split the diagonal of the listed rows into two entries (d/2 + d/2)
        rows, cols, vals = [], [], []
        for r in range(A.shape[0]):
            for o in range(rp[r], rp[r + 1]):
                if ci[o] == r and r in dup_diag:
                    rows += [r, r]
                    cols += [r, r]
                    vals += [va[o] / 2, va[o] / 2]
                else:
                    rows.append(r)
                    cols.append(ci[o])
                    vals.append(va[o])
        rp = np.zeros(A.shape[0] + 1, np.int32)
        for r in rows:
            rp[r + 1] += 1
        rp = np.cumsum(rp).astype(np.int32)
        ci, va = np.asarray(cols, np.int32), np.asarray(vals)
    return X.Csr(rp, ci, va, np.ones(A.shape[0]), np.zeros(A.shape[0]))


def dominant(n, seed, density=0.1, margin=1.5):
    rng = np.random.default_rng(seed)
    M = (rng.random((n, n)) < density) * rng.uniform(-1, 1, (n, n))
    np.fill_diagonal(M, 0.0)
    d = np.abs(M).sum(1) * margin + 0.5
    M[np.arange(n), np.arange(n)] = d * np.where(rng.random(n) < 0.5, 1, -1)
    return M


def f32(v):
    return np.asarray(v, np.float64).astype(np.float32).astype(np.float64)


def test_diagonal_system_one_update():
    d = np.array([2.0, -3.0, 0.7, 1e-3, 5.0])
    b = np.array([1.0, 1.0, -2.5, 3.0, 0.1])
    a = csr_of(np.diag(d))
    r = O.jacobi(a, b=b, max_iters=1)
    want = (np.float32(1.0) / d.astype(np.float32)) * b.astype(np.float32)
    np.testing.assert_array_equal(r["x"], want)
    assert r["iterations"] == 1 and r["status"] == 1
    assert r["diff"] == np.abs(want).max()
    r3 = O.jacobi(a, b=b, max_iters=3)
    np.testing.assert_array_equal(r3["x"], want)
    assert r3["diff"] == 0.0


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_error_recursion(seed):
    n = 40
    A = dominant(n, seed, margin=1.1)
    a = csr_of(A)
    rng = np.random.default_rng(seed + 10)
    b = rng.uniform(-1, 1, n)
    x0 = rng.uniform(-1, 1, n)
    A32 = f32(A)
    D = np.diag(np.diag(A32))
    R = A32 - D
    B = -np.linalg.solve(D, R)
    xs = np.linalg.solve(A32, f32(b))
    e = f32(x0) - xs
    for k in range(1, 9):
        r = O.jacobi(a, b=b, x0=x0, max_iters=k)
        assert r["iterations"] == k
        e = B @ e
        np.testing.assert_allclose(r["x"], xs + e, rtol=0, atol=2e-5 * max(1.0, np.abs(xs).max()))


def test_fixed_point_and_tau_stop():
    n = 200
    A = dominant(n, 7, density=0.05, margin=2.0)
    a = csr_of(A)
    b = np.random.default_rng(3).uniform(-1, 1, n)
    xs = np.linalg.solve(A, b)
    r = O.jacobi(a, b=b, max_iters=200)
    assert np.abs(r["x"] - xs).max() <= 1e-5 * np.abs(xs).max()
    t = O.jacobi(a, b=b, max_iters=10000, use_tau=True, tau=1e-6)
    assert t["status"] == 0 and t["diff"] <= 1e-6 < O.jacobi(a, b=b, max_iters=t["iterations"] - 1)["diff"]
    # the fixed-count run of the same length gives the same vector
    np.testing.assert_array_equal(t["x"], O.jacobi(a, b=b, max_iters=t["iterations"])["x"])


def test_diff_max_brute_force():
    n = 30
    A = dominant(n, 11)
    a = csr_of(A)
    b = np.linspace(-1, 1, n)
    prev = O.jacobi(a, b=b, max_iters=0)["x"]
    for k in range(1, 6):
        r = O.jacobi(a, b=b, max_iters=k)
        assert r["diff"] == np.abs(r["x"] - prev).max()
        prev = r["x"]


def test_duplicate_diagonal_is_summed():
    n = 25
    A = dominant(n, 5)
    A = np.round(A * 64) / 64            # dyadic: d/2 + d/2 == d exactly
    b = np.random.default_rng(1).uniform(-1, 1, n)
    r1 = O.jacobi(csr_of(A), b=b, max_iters=5)
    r2 = O.jacobi(csr_of(A, dup_diag={0, 3, 24}), b=b, max_iters=5)
    np.testing.assert_array_equal(r1["x"], r2["x"])


def test_bad_diagonal():
    A = np.diag([1.0, 2.0, 3.0])
    A[1, 1] = 0.0
    A[1, 0] = 1.0
    with pytest.raises(O.OracleError, match="missing or zero diagonal at row 1"):
        O.jacobi(csr_of(A), b=np.ones(3))
    # a diagonal that cancels to zero
    a = X.Csr(np.array([0, 1, 3], np.int32), np.array([0, 1, 1], np.int32), np.array([1.0, 2.0, -2.0]))
    with pytest.raises(O.OracleError, match="missing or zero diagonal at row 1"):
        O.jacobi(a, b=np.ones(2))


def test_ones_rhs_matches_row_sums():
    rng = np.random.default_rng(9)
    A = dominant(50, 9)
    A = np.round(A * 8)
    np.fill_diagonal(A, np.abs(A).sum(1) + 1)
    a = csr_of(A)
    r = O.jacobi(a, b=None, max_iters=0)
    np.testing.assert_array_equal(r["b"], A.sum(1).astype(np.float32))
    np.testing.assert_array_equal(X.rhs_ones_f32(a), r["b"].astype(np.float64))
    # non-integer values: library and oracle agree bit for bit
    a = X.gen_fem2d(20, 1.0, 2, 1e3, 0.3, 4)
    np.testing.assert_array_equal(X.rhs_ones_f32(a), O.jacobi_prepare(a)["b"].astype(np.float64))
    del rng


def test_tiny_case_float32_scalars():
    """numpy float32 scalar arithmetic in the reference's loop order."""
    A = np.array([[4.0, -1.3, 0.0, 0.2], [-1.3, 5.1, 0.7, 0.0], [0.0, 0.7, 3.3, -0.9], [0.2, 0.0, -0.9, 2.8]])
    b = np.array([1.1, -0.4, 2.2, 0.3])
    a = csr_of(A)
    x = [np.float32(0)] * 4
    for _ in range(6):
        nx = []
        for r in range(4):
            rx = np.float32(0)
            for o in range(a.row_ptr[r], a.row_ptr[r + 1]):
                c = a.col_idx[o]
                if c != r:
                    rx = np.float32(rx + np.float32(x[c] * np.float32(a.values[o])))
            d = np.float32(a.values[a.row_ptr[r] + list(a.col_idx[a.row_ptr[r]:a.row_ptr[r + 1]]).index(r)])
            nx.append(np.float32(np.float32(np.float32(b[r]) - rx) * np.float32(np.float32(1) / d)))
        x = nx
    np.testing.assert_array_equal(O.jacobi(a, b=b, max_iters=6)["x"], np.array(x, np.float32))


def test_fixed_count_runs_through_breakdown():
    """Without use_tau every update runs, as the reference's RunJacobiCpu;
    the status says breakdown when the last difference is not finite.  With
    use_tau the first infinite difference stops the loop."""
    A = np.kron(np.eye(8), np.array([[1.0, 2.0], [2.0, 1.0]]))
    a = csr_of(A)
    r = O.jacobi(a, b=np.ones(16), max_iters=400)
    assert r["iterations"] == 400 and r["status"] == 2 and not np.isfinite(r["diff"])
    t = O.jacobi(a, b=np.ones(16), max_iters=400, use_tau=True, tau=1e-30)
    assert t["status"] == 2 and t["iterations"] < 400 and np.isinf(t["diff"])
    # the updates before the overflow are the same in both forms
    u = O.jacobi(a, b=np.ones(16), max_iters=t["iterations"])
    np.testing.assert_array_equal(u["x"], t["x"])

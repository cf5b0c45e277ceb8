"""GPU plain Jacobi iteration (xp_jacobi_*) against the FP32 oracle
(orc_jacobi_prepare_f32 + orc_jacobi_run_f32, the reference's
PrepareJacobiVectorsAndR + RunJacobiCpu).

Bar: bit for bit.  The kernel sums each row's R x in CSR order with one
rounding per product and per sum, as the reference's CPU loop does, so x
(every entry), the update count, the status and the final diff_max must be
identical -- on single systems, batches, duplicate and unsorted entries,
early stops (use_tau), breakdown and the graph / direct launch paths, in
both forms: the fixed count (k_jacobi_sweep, no stop test) and the
checked iteration (k_jacobi_iter, use_tau).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle_lib as O
import xplus_amd as X

pytestmark = pytest.mark.gpu


def dominant(n, seed, density=0.02, margin=1.2):
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    M = sp.random(n, n, density=density, random_state=rng, format="lil")
    M = sp.csr_matrix(M)
    M.data = rng.uniform(-1, 1, M.nnz)
    M.setdiag(0.0)
    M.eliminate_zeros()
    d = np.asarray(abs(M).sum(1)).ravel() * margin + 0.5
    M = M + sp.diags(d * np.where(rng.random(n) < 0.5, 1, -1))
    M = sp.csr_matrix(M)
    M.sort_indices()
    return X.Csr(M.indptr, M.indices, M.data, rng.uniform(-1, 1, n), np.zeros(n))


def shuffled_with_duplicates(a, seed):
    """Same matrix: columns shuffled inside each row, some entries (the
    diagonal among them) split into two exact halves."""
    rng = np.random.default_rng(seed)
    rp, ci, va = [0], [], []
    for r in range(a.n):
        cols = list(a.col_idx[a.row_ptr[r]:a.row_ptr[r + 1]])
        vals = list(a.values[a.row_ptr[r]:a.row_ptr[r + 1]])
        ent = []
        for c, v in zip(cols, vals):
            if rng.random() < 0.3:
                ent += [(c, v / 2), (c, v / 2)]
            else:
                ent.append((c, v))
        rng.shuffle(ent)
        ci += [c for c, _ in ent]
        va += [v for _, v in ent]
        rp.append(len(ci))
    return X.Csr(np.asarray(rp), np.asarray(ci), np.asarray(va), a.b, a.x0)


def check(a_list, b_list=None, x0_list=None, max_iters=25, use_tau=False, tau=1e-5, use_graph=True, graph_iters=0,
          **mkw):
    m = X.Matrix(a_list if len(a_list) > 1 else a_list[0], **mkw)
    s = X.Jacobi(m, max_iters=max_iters, tau=tau, use_tau=use_tau, use_graph=use_graph, graph_iters=graph_iters)
    bs = b_list or [a.b for a in a_list]
    x0s = x0_list or [None] * len(a_list)
    x0cat = None if all(v is None for v in x0s) else np.concatenate(
        [np.zeros(a.n) if v is None else v for a, v in zip(a_list, x0s)])
    s.set_rhs(np.concatenate(bs), x0cat)
    s.run()
    x = s.x()
    off = 0
    out = []
    for k, a in enumerate(a_list):
        ref = O.jacobi(a, b=bs[k], x0=x0s[k], max_iters=max_iters, use_tau=use_tau, tau=tau)
        st = s.status(k)
        print(f"system {k}: n={a.n} iters {st['iterations']} (oracle {ref['iterations']}) status {st['status']} "
              f"diff {st['diff']:.3e}", flush=True)
        assert st["iterations"] == ref["iterations"]
        assert st["status"] == ref["status"]
        assert np.array_equal(np.float32(st["diff"]), np.float32(ref["diff"]), equal_nan=True)
        assert np.array_equal(x[off:off + a.n], ref["x"].astype(np.float64), equal_nan=True)
        off += a.n
        out.append((st, ref))
    return s, out


@pytest.mark.parametrize("use_graph", [True, False])
def test_fixed_count_graph_remainder(gpu, use_graph):
    """A count that is not a multiple of the graph length: whole graphs,
    then the remainder launched directly, with launch parity carried over."""
    check([dominant(2500, 14)], max_iters=23, graph_iters=6, use_graph=use_graph)


@pytest.mark.parametrize("which", ["c1", "diff3d", "dominant", "dup", "big"])
def test_fixed_count_bit_exact(gpu, which):
    a = {"c1": lambda: X.gen_diffusion2d(32, 32, 1.0, 20261204),
         "diff3d": lambda: X.gen_diffusion3d(16, 1.5, 3),
         "dominant": lambda: dominant(3000, 4),
         "dup": lambda: shuffled_with_duplicates(dominant(2000, 6), 7),
         "big": lambda: X.gen_diffusion2d(512, 512, 1.0, 9)}[which]()
    print(f"jacobi {which}: n={a.n} nnz={a.nnz}", flush=True)
    check([a], max_iters=40 if which != "big" else 12)


def test_reference_rhs_and_start(gpu):
    """b = A * ones (the reference's rule without b) and a nonzero start."""
    a = dominant(1500, 12)
    b = X.rhs_ones_f32(a)
    np.testing.assert_array_equal(b, O.jacobi_prepare(a)["b"].astype(np.float64))
    x0 = np.random.default_rng(1).uniform(-2, 2, a.n)
    check([a], b_list=[b], x0_list=[x0], max_iters=17)


@pytest.mark.parametrize("use_graph", [True, False])
def test_tau_stop(gpu, use_graph):
    a = dominant(4000, 21, margin=1.5)
    _, out = check([a], max_iters=100000, use_tau=True, tau=1e-6, use_graph=use_graph, graph_iters=6)
    st, _ = out[0]
    assert st["status"] == 0 and st["diff"] <= 1e-6


def test_batch_uneven(gpu):
    systems = [X.gen_diffusion2d(20, 30, 1.0, 1), dominant(700, 2), shuffled_with_duplicates(dominant(333, 3), 4)]
    check(systems, max_iters=31)
    check(systems, max_iters=100000, use_tau=True, tau=1e-5)


def test_zero_iterations_and_repeat_runs(gpu):
    a = dominant(500, 8)
    x0 = np.linspace(-1, 1, a.n)
    s, out = check([a], x0_list=[x0], max_iters=0)
    assert out[0][0]["iterations"] == 0 and out[0][0]["status"] == 1
    np.testing.assert_array_equal(s.x(), x0.astype(np.float32).astype(np.float64))
    # a second run starts again from x0
    s2, _ = check([a], max_iters=9)
    s2.run()
    np.testing.assert_array_equal(s2.x(), O.jacobi(a, b=a.b, max_iters=9)["x"].astype(np.float64))


@pytest.mark.parametrize("use_tau", [False, True])
def test_breakdown_on_divergence(gpu, use_tau):
    """rho(D^-1 R) = 2: the iterates double until they overflow.  Fixed
    count: all updates run (as the reference's), the status is breakdown as
    the last difference is not finite.  With the stop test the first
    infinite difference stops the run."""
    import scipy.sparse as sp
    n = 256
    blk = sp.csr_matrix(np.array([[1.0, 2.0], [2.0, 1.0]]))
    A = sp.csr_matrix(sp.block_diag([blk] * (n // 2)))
    a = X.Csr(A.indptr, A.indices, A.data, np.ones(n), np.zeros(n))
    _, out = check([a], max_iters=1000, use_tau=use_tau, tau=1e-30)
    assert out[0][0]["status"] == 2
    assert (out[0][0]["iterations"] == 1000) != use_tau


def test_missing_diagonal_rejected(gpu):
    a = X.Csr(np.array([0, 1, 2, 4]), np.array([0, 0, 1, 2]), np.array([1.0, 1.0, 1.0, 1.0]))
    m = X.Matrix(a)
    with pytest.raises(X.XplusError, match="missing or zero diagonal at row 1"):
        X.Jacobi(m)


def test_ingested_mtx(gpu, tmp_path):
    a0 = dominant(1200, 30)
    import scipy.sparse as sp
    A = a0.to_scipy().tocoo()
    lines = [f"{r + 1} {c + 1} {float(v)!r}" for r, c, v in zip(A.row, A.col, A.data)]
    (tmp_path / "d.mtx").write_text("%%MatrixMarket matrix coordinate real general\n"
                                    f"{a0.n} {a0.n} {len(lines)}\n" + "\n".join(lines) + "\n")
    a, _ = X.load_mtx(tmp_path / "d.mtx")
    del sp
    check([a], max_iters=20)


def test_kernel_time(gpu):
    a = X.gen_diffusion2d(256, 256, 1.0, 2)
    m = X.Matrix(a)
    s = X.Jacobi(m, max_iters=10)
    s.set_rhs(a.b)
    us = s.kernel_time(10)
    print(f"k_jacobi_iter 256^2: {us:.2f} us", flush=True)
    assert us > 0.0

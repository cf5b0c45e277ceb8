"""Strength-aggregated 8-row blocks: the oracle (oracle/blocking.py) pinned by
invariants and closed-form cases, the library's C++ aggregation
(xp_strength_blocking, no GPU) against it, and the block features over
aggregates against natural-group features where the two coincide."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle_lib as O
import xplus_amd as X

sys.path.insert(0, str(O.ROOT / "oracle"))
import blocking as BK  # noqa: E402
import block_mlp as B  # noqa: E402


def csr(A):
    import scipy.sparse as sp
    A = sp.csr_matrix(A)
    A.sort_indices()
    return X.Csr(A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64),
                 np.ones(A.shape[0]), np.zeros(A.shape[0]))


def random_spd(n, seed, density=0.08):
    rng = np.random.default_rng(seed)
    M = (rng.random((n, n)) < density) * rng.standard_normal((n, n))
    M = np.triu(M, 1)
    M = M + M.T
    np.fill_diagonal(M, np.abs(M).sum(1) + 1.0 + rng.random(n))
    return M


def check_partition(blocks, n):
    rows = [r for b in blocks for r in b]
    assert sorted(rows) == list(range(n))
    assert all(1 <= len(b) <= 8 for b in blocks)
    assert all(b == sorted(b) for b in blocks)
    assert [b[0] for b in blocks] == sorted(b[0] for b in blocks)


def test_oracle_chain_gives_consecutive_runs():
    """A path graph (tridiagonal matrix, any positive couplings): every
    block is a run of consecutive rows; with n = 8 k + 5 the runs of 8 and
    a last run of 5."""
    n = 37
    rng = np.random.default_rng(1)
    w = rng.uniform(0.5, 2.0, n - 1)
    A = np.diag(np.r_[w, 0] + np.r_[0, w] + 0.1) - np.diag(w, 1) - np.diag(w, -1)
    a = csr(A)
    blocks = BK.strength_blocking(a.row_ptr, a.col_idx, a.values)
    check_partition(blocks, n)
    assert blocks == [list(range(i, min(i + 8, n))) for i in range(0, n, 8)]


def test_oracle_uniform_grid_gives_grid_rows():
    """5-point Laplacian on an 8 x 8 grid, uniform strengths: ties go to the
    smaller row, so every block is one grid row."""
    a = X.gen_diffusion2d(8, 8, 0.0, 3)    # sigma = 0: all conductivities 1
    blocks = BK.strength_blocking(a.row_ptr, a.col_idx, a.values)
    assert blocks == [list(range(8 * i, 8 * i + 8)) for i in range(8)]


def test_oracle_first_block_is_greedy():
    """Brute force on the first block of a random matrix: each added row has
    the largest summed strength to the rows already in the block."""
    A = random_spd(60, 7)
    a = csr(A)
    s = np.abs(A) / np.sqrt(np.outer(np.diag(A), np.diag(A)))
    np.fill_diagonal(s, 0.0)
    blocks = BK.strength_blocking(a.row_ptr, a.col_idx, a.values)
    check_partition(blocks, 60)
    # replay the growth of the block seeded at row 0 (first pass, nothing assigned yet)
    blk = [0]
    while len(blk) < 8:
        w = s[blk].sum(0)
        w[blk] = 0.0
        if w.max() <= 0.0:
            break
        j = int(np.flatnonzero(w == w.max())[0])
        blk.append(j)
    if len(blk) == 8:                      # a full first block survives both passes unchanged
        assert sorted(blk) in blocks


@pytest.mark.parametrize("case", ["c1", "fem", "random", "chain_ragged"])
def test_library_blocking_matches_oracle(case):
    if case == "c1":
        a = X.gen_diffusion2d(32, 32, 1.0, 20265001)
    elif case == "fem":
        a = X.gen_fem2d(34, 1.0, 4, 1e3, 0.3, 5)
    elif case == "random":
        a = csr(random_spd(300, 11))
    else:
        a = X.gen_diffusion2d(13, 11, 1.0, 7)
    ref = BK.strength_blocking(a.row_ptr, a.col_idx, a.values)
    got = X.strength_blocking(a)
    got = [[int(r) for r in row if r >= 0] for row in got]
    check_partition(got, a.n)
    assert got == ref


def test_group_features_reduce_to_natural():
    a = X.gen_fem2d(20, 1.0, 2, 1e3, 0.3, 9)
    f1 = B.features_from_csr(a.row_ptr, a.col_idx, a.values)
    f2 = B.features_from_csr(a.row_ptr, a.col_idx, a.values, B.natural_groups(a.n))
    for u, v in zip(f1, f2):
        np.testing.assert_array_equal(u, v)


def test_library_blockset_over_aggregates_matches_oracle():
    a = X.gen_fem2d(26, 1.0, 3, 1e3, 0.3, 13)
    groups = BK.strength_blocking(a.row_ptr, a.col_idx, a.values)
    ref = B.features_from_csr(a.row_ptr, a.col_idx, a.values, groups)
    got = X.BlockSet(a, blocking=1).arrays()
    for name, g, r in zip(("feat", "ahat", "hmat", "scale"), got, ref):
        np.testing.assert_allclose(g, r, rtol=1e-6, atol=1e-6 * max(1.0, np.abs(r).max()), err_msg=name)


def test_aggregates_beat_strips_with_exact_blocks():
    """FP64 oracle block PCG with exact block inverses: aggregates need fewer
    iterations than 8 consecutive rows on the FEM family."""
    a = X.gen_fem2d(66, 1.0, 6, 1e3, 0.3, 20263203)
    A = a.to_scipy().tocsr()
    tau = (1e-6 * np.linalg.norm(a.b)) ** 2
    its = {}
    for name, groups in (("strips", B.natural_groups(a.n)),
                         ("aggregates", BK.strength_blocking(a.row_ptr, a.col_idx, a.values))):
        blocks = np.zeros((len(groups), 8, 8))
        for k, g in enumerate(groups):
            m = len(g)
            blocks[k, :m, :m] = np.linalg.inv(A[g][:, g].toarray())
        _, it, _ = B.block_pcg(A, a.b, blocks, tau, groups=groups)
        its[name] = it
    print(its)
    assert its["aggregates"] < 0.9 * its["strips"]

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


@pytest.mark.parametrize("case", ["c1", "fem", "random", "chain_ragged"])
def test_library_ordered_blocking_matches_oracle(case):
    if case == "c1":
        a = X.gen_diffusion2d(32, 32, 1.0, 20265001)
    elif case == "fem":
        a = X.gen_fem2d(34, 1.0, 4, 1e3, 0.3, 5)
    elif case == "random":
        a = csr(random_spd(300, 11))
    else:
        a = X.gen_diffusion2d(13, 11, 1.0, 7)
    ref = BK.strength_blocking(a.row_ptr, a.col_idx, a.values, rule=2)
    got = X.strength_blocking(a, rule=2)
    got = [[int(r) for r in row if r >= 0] for row in got]
    check_partition(got, a.n)
    assert got == ref
    assert ref != BK.strength_blocking(a.row_ptr, a.col_idx, a.values)


@pytest.mark.parametrize("seed", range(6))
def test_library_ordered_blocking_matches_oracle_random(seed):
    """Rule 2 on random matrices of assorted sizes and densities, including
    exact ties in the strengths (integer values): library == oracle."""
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(9, 400))
    A = random_spd(n, 200 + seed, density=float(rng.uniform(0.01, 0.1)))
    if seed % 2:
        A = np.round(A)                     # many equal strengths: the tie rules decide
        np.fill_diagonal(A, np.abs(A).sum(1) + 1.0)
    a = csr(A)
    ref = BK.strength_blocking(a.row_ptr, a.col_idx, a.values, rule=2)
    got = [[int(r) for r in row if r >= 0] for row in X.strength_blocking(a, rule=2)]
    check_partition(got, n)
    assert got == ref


def test_oracle_ordered_first_block_brute_force():
    """Rule 2 replayed with dense numpy: the first seed is the row of largest
    total strength, and each added row has the largest fraction of its total
    strength going to the rows already in the block."""
    for seed in (7, 8, 9):
        A = random_spd(60, seed)
        a = csr(A)
        s = np.abs(A) / np.sqrt(np.outer(np.diag(A), np.diag(A)))
        np.fill_diagonal(s, 0.0)
        t = s.sum(1)
        blocks = BK.strength_blocking(a.row_ptr, a.col_idx, a.values, rule=2)
        check_partition(blocks, 60)
        blk = [int(np.argmax(t))]
        while len(blk) < 8:
            w = s[blk].sum(0) / t
            w[blk] = -1.0
            w[s[blk].sum(0) == 0.0] = -1.0         # only neighbours of the block are candidates
            if w.max() < 0.0:
                break
            blk.append(int(np.flatnonzero(w == w.max())[0]))
        if len(blk) == 8:
            assert sorted(blk) in blocks


def test_oracle_ordered_chain_keeps_strong_edge():
    """Hand-derived case.  Path graph, n = 16, unit couplings except 50 on
    edge 7-8, diagonal = couplings + 1.  Strengths: interior 1/3, ends
    s01 = 1/sqrt(6), strong edge 50/52, its neighbours 1/sqrt(156); totals
    t7 = t8 = 1.04 (largest), t1 = t14 = 0.741, t0 = 0.408.  Rule 2 seeds row
    7, takes 8 (fraction 0.92), then 6 (0.19, tie with 9 -> smaller), then
    5, 4, 3, 2 (0.5 each) and 1 (0.45 > 0.19): block 1..8.  Seed 14 grows
    15, 13 .. 9 (7 rows, no candidate left), seed 0 stays alone; pass 2
    regroups the same two pieces and packing merges them.  Rule 1 cuts the
    strong edge (runs 0..7, 8..15)."""
    n = 16
    w = np.ones(n - 1)
    w[7] = 50.0
    A = np.diag(np.r_[w, 0] + np.r_[0, w] + 1.0) - np.diag(w, 1) - np.diag(w, -1)
    a = csr(A)
    got = BK.strength_blocking(a.row_ptr, a.col_idx, a.values, rule=2)
    assert got == [[0, 9, 10, 11, 12, 13, 14, 15], list(range(1, 9))]
    assert BK.strength_blocking(a.row_ptr, a.col_idx, a.values) == [list(range(8)), list(range(8, 16))]
    lib = [[int(r) for r in row if r >= 0] for row in X.strength_blocking(a, rule=2)]
    assert lib == got


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
                         ("aggregates", BK.strength_blocking(a.row_ptr, a.col_idx, a.values)),
                         ("ordered", BK.strength_blocking(a.row_ptr, a.col_idx, a.values, rule=2))):
        blocks = np.zeros((len(groups), 8, 8))
        for k, g in enumerate(groups):
            m = len(g)
            blocks[k, :m, :m] = np.linalg.inv(A[g][:, g].toarray())
        _, it, _ = B.block_pcg(A, a.b, blocks, tau, groups=groups)
        its[name] = it
    print(its)
    assert its["aggregates"] < 0.9 * its["strips"]
    assert its["ordered"] < its["aggregates"]


# ---------------------------------------------------------------- rule 3 (device aggregation)
def test_parallel_rule_partitions():
    for a in (csr(random_spd(90, 11)), X.gen_fem2d(26, 1.0, 3, 1e3, 0.3, 5), X.gen_diffusion2d(13, 11, 1.0, 7)):
        blocks = BK.parallel_blocking(list(a.row_ptr), list(a.col_idx), list(a.values))
        check_partition(blocks, a.n)


def test_parallel_rule_recovers_cliques():
    """Disjoint 8-row cliques under a random row permutation: within distance
    2 of a row there are only rows of its clique, so every clique yields one
    seed and grows into exactly that clique."""
    rng = np.random.default_rng(4)
    k = 9
    A = np.zeros((8 * k, 8 * k))
    for c in range(k):
        blk = rng.uniform(0.2, 1.0, (8, 8))
        blk = blk + blk.T
        np.fill_diagonal(blk, blk.sum(1) + 1.0)
        A[8 * c:8 * c + 8, 8 * c:8 * c + 8] = blk
    perm = rng.permutation(8 * k)
    P = A[np.ix_(perm, perm)]
    a = csr(P)
    blocks = BK.parallel_blocking(list(a.row_ptr), list(a.col_idx), list(a.values))
    inv = np.argsort(perm)
    want = sorted(sorted(int(inv[8 * c + j]) for j in range(8)) for c in range(k))
    assert sorted(blocks) == want


def test_parallel_rule_seeds_brute_force():
    """Every round's seeds: the free rows that precede every free row they
    reach in at most two steps along the rows' entries (brute force)."""
    A = random_spd(70, 12, density=0.06)
    a = csr(A)
    rp, ci = list(a.row_ptr), list(a.col_idx)
    s = BK.strengths(rp, ci, list(a.values))
    t = BK.totals(rp, ci, s)
    rng = np.random.default_rng(3)
    for trial in range(5):
        owner = [int(v) for v in np.where(rng.random(70) < 0.3 * trial / 4, 0, -1)]
        elig = [True] * 70
        nb = {i: set(ci[rp[i]:rp[i + 1]]) - {i} for i in range(70)}
        want = []
        for i in range(70):
            if owner[i] >= 0:
                continue
            reach = {i} | nb[i] | {d for c in nb[i] for d in nb[c]}
            rivals = [j for j in reach if owner[j] < 0 and j != i]
            if all(t[i] > t[j] or (t[i] == t[j] and i < j) for j in rivals):
                want.append(i)
        assert BK.seeds_of_round(rp, ci, t, elig, owner) == want


def test_rule4_seeds_brute_force():
    """Rule 4 (seed_dist 1): every round's seeds are the free rows that
    precede every free row they reach in one step (brute force); on the same
    state they include every rule-3 seed."""
    A = random_spd(70, 13, density=0.08)
    a = csr(A)
    rp, ci = list(a.row_ptr), list(a.col_idx)
    s = BK.strengths(rp, ci, list(a.values))
    t = BK.totals(rp, ci, s)
    rng = np.random.default_rng(5)
    for trial in range(5):
        owner = [int(v) for v in np.where(rng.random(70) < 0.3 * trial / 4, 0, -1)]
        elig = [True] * 70
        nb = {i: set(ci[rp[i]:rp[i + 1]]) - {i} for i in range(70)}
        want = []
        for i in range(70):
            if owner[i] >= 0:
                continue
            rivals = [j for j in nb[i] if owner[j] < 0]
            if all(t[i] > t[j] or (t[i] == t[j] and i < j) for j in rivals):
                want.append(i)
        got = BK.seeds_of_round(rp, ci, t, elig, owner, seed_dist=1)
        assert got == want
        assert set(BK.seeds_of_round(rp, ci, t, elig, owner)) <= set(got)


def test_rule4_partitions_and_takes_fewer_rounds_on_dense_rows():
    """Rule 4 still partitions the rows into blocks of at most 8; on a dense
    matrix its first round seeds more blocks than rule 3's."""
    for a in (csr(random_spd(90, 11)), X.gen_fem2d(26, 1.0, 3, 1e3, 0.3, 5), csr(random_spd(120, 17, density=0.25))):
        check_partition(BK.parallel_blocking(list(a.row_ptr), list(a.col_idx), list(a.values), seed_dist=1), a.n)
    a = csr(random_spd(120, 17, density=0.25))
    rp, ci = list(a.row_ptr), list(a.col_idx)
    t = BK.totals(rp, ci, BK.strengths(rp, ci, list(a.values)))
    free = [-1] * a.n
    n1 = len(BK.seeds_of_round(rp, ci, t, [True] * a.n, free, seed_dist=1))
    n2 = len(BK.seeds_of_round(rp, ci, t, [True] * a.n, free))
    assert n1 > n2


def test_parallel_rule_growth_steps_replayed():
    """Every growth step of a full run, replayed from the owner map it chose
    from: each taken row bid for its block with the best score among the
    blocks it touches, and no row with a larger FP32-rounded score bid for
    the same block."""
    A = random_spd(80, 13, density=0.07)
    a = csr(A)
    rp, ci, va = list(a.row_ptr), list(a.col_idx), list(a.values)
    s = BK.strengths(rp, ci, va)
    t = BK.totals(rp, ci, s)
    log = []
    BK.parallel_blocking(rp, ci, va, log=log)
    S = np.zeros((80, 80))
    for r in range(80):
        for o in range(rp[r], rp[r + 1]):
            if ci[o] != r:
                S[r, ci[o]] += s[o]
    checked = 0
    for prev, take in log:
        if take:
            size = {}
            for r, B in enumerate(prev):
                if B >= 0:
                    size[B] = size.get(B, 0) + 1
            score = {}
            for j in range(80):
                if prev[j] >= 0:
                    continue
                fr = {}
                for B in set(prev[c] for c in range(80) if S[j, c] > 0 and prev[c] >= 0):
                    if size[B] < 8:
                        fr[B] = sum(S[j, c] for c in range(80) if prev[c] == B) / t[j]
                if fr:
                    score[j] = max(fr.items(), key=lambda kv: (kv[1], -kv[0]))
            for B, j in take.items():
                assert score[j][0] == B
                rivals = [k for k, (b2, f2) in score.items() if b2 == B]
                assert all(np.float32(score[k][1]) <= np.float32(score[j][1]) + 0.0 for k in rivals)
                checked += 1
    assert checked > 20


def test_parallel_rule_matches_ordered_rule_quality():
    """Rule 3 exists to run on the device; it must keep rule 2's blocks'
    quality: exact block inverses over its aggregates need no more than 3%
    more PCG iterations than over rule 2's (FP64, mean of three right-hand
    sides)."""
    a = X.gen_fem2d(66, 1.0, 6, 1e3, 0.3, 20265600)
    A = a.to_scipy()
    its = {}
    for rule in (2, 3):
        groups = (BK.parallel_blocking(list(a.row_ptr), list(a.col_idx), list(a.values)) if rule == 3
                  else BK.strength_blocking(a.row_ptr, a.col_idx, a.values, rule=2))
        blocks = np.zeros((len(groups), 8, 8))
        for b, g in enumerate(groups):
            blocks[b, :len(g), :len(g)] = np.linalg.inv(A[g][:, g].toarray())
            blocks[b, len(g):, len(g):] = np.eye(8 - len(g))
        n_it = []
        for k in range(3):
            rhs = np.random.default_rng(k).uniform(-1, 1, a.n)
            n_it.append(B.block_pcg(A, rhs, blocks, (1e-6 * np.linalg.norm(rhs)) ** 2, groups=groups)[1])
        its[rule] = float(np.mean(n_it))
    print("exact-block iterations by rule:", its)
    assert its[3] <= 1.03 * its[2]

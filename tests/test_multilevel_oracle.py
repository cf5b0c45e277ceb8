"""Pins of the multilevel oracle (oracle/multilevel.py), on the CPU.

  * plain_aggregation meets its definition, checked by breadth-first search
    on the strength graph: roots pairwise at strong distance >= 3, every
    row within distance 2 of a root (maximality), a root's strong
    neighbours in its aggregate, every other row in the aggregate of its
    strongest joined neighbour, root k = aggregate k in index order; and it
    equals a brute-force restatement of the rounds (all-pairs distances
    instead of max propagation) on small graphs, rounds included;
  * galerkin equals scipy's P^T A P (an independent formula: sparse
    products) to FP64 rounding, and a hand-computed 6-row case bit for bit
    (the order of the sums is part of the definition);
  * vcycle on two levels equals the cycle written as one dense matrix,
    V = (I - B A)(B + C (I - A B)) + B with B = omega M, C = gamma P Ac^-1 P^T;
  * the cycle's matrix (three levels) is symmetric and positive definite,
    which a PCG preconditioner must be, at the parameters the solver uses;
  * PCG with the cycle converges to numpy's solution of the system.
"""
from __future__ import annotations

import sys
from collections import deque

import numpy as np
import pytest
import scipy.sparse as sp

import oracle_lib as O
import xplus_amd as X

sys.path.insert(0, str(O.ROOT / "oracle"))
import blocking as BK  # noqa: E402
import multilevel as ML  # noqa: E402

THETA = 0.08


def hierarchy(a, nlev, omega=0.7, gamma=1.5, seed_dist=2, theta=THETA):
    """Levels: smoother blocks from the rule-3 (then rule seed_dist)
    aggregation oracle with exact FP64 inverses, coarse spaces from plain
    aggregation; dense coarsest."""
    A = a.to_scipy().tocsr()
    rp, ci, va = A.indptr, A.indices, A.data
    levels = []
    for lev in range(nlev):
        groups = BK.parallel_blocking(list(rp), list(ci), list(va), seed_dist=2 if lev == 0 else seed_dist)
        slots = np.full((len(groups), 8), -1, np.int32)
        for k, g in enumerate(groups):
            slots[k, :len(g)] = g
        Am = sp.csr_matrix((va, ci, rp), shape=(len(rp) - 1, len(rp) - 1))
        agg, nagg, _ = ML.plain_aggregation(rp, ci, va, theta, lev)
        levels.append(ML.Level(Am, slots, ML.exact_blocks(Am, slots, f32_values=False), agg, nagg))
        rp, ci, va = ML.galerkin(rp, ci, va, agg, nagg)
    coarse = ML.to_dense(rp, ci, va, len(rp) - 1)
    return ML.Hierarchy(levels, coarse, omega, gamma)


def bfs_dist(adj, src, cap=3):
    """Strong distances from src up to cap (cap for farther rows)."""
    d = {src: 0}
    q = deque([src])
    while q:
        u = q.popleft()
        if d[u] >= cap:
            continue
        for v in adj[u]:
            if v not in d:
                d[v] = d[u] + 1
                q.append(v)
    return d


def adjacency(rp, ci, va, theta):
    sp_, si, sv = ML.strong_graph(rp, ci, va, theta)
    n = len(rp) - 1
    adj = [list(si[sp_[i]:sp_[i + 1]]) for i in range(n)]
    w = [dict(zip(si[sp_[i]:sp_[i + 1]], sv[sp_[i]:sp_[i + 1]])) for i in range(n)]
    return adj, w


def check_aggregation_properties(rp, ci, va, theta, salt):
    agg, nagg, rounds, roots = ML.plain_aggregation(rp, ci, va, theta, salt, with_roots=True)
    n = len(rp) - 1
    adj, w = adjacency(rp, ci, va, theta)
    assert agg.shape == (n,) and agg.min() >= 0 and agg.max() == nagg - 1
    assert set(agg.tolist()) == set(range(nagg))
    roots = [int(r) for r in roots]
    assert roots == sorted(roots)
    dist = [bfs_dist(adj, i) for i in range(n)]
    # root k is aggregate k; its strong neighbours share it; roots pairwise
    # at distance >= 3
    for k, r in enumerate(roots):
        assert agg[r] == k
        assert all(agg[j] == k for j in adj[r])
    for x in roots:
        for y in roots:
            if x < y:
                assert dist[x].get(y, 3) >= 3
    rs = set(roots)
    # maximality: every row within distance 2 of a root
    for i in range(n):
        assert any(dist[i].get(r, 3) <= 2 for r in roots), i
    # the other rows: the aggregate of the strongest neighbour joined to a
    # root (largest |a_ij|, then the smallest j); none are left over here
    joined = {i for i in range(n) if i in rs or any(j in rs for j in adj[i])}
    for i in range(n):
        if i in joined:
            continue
        cand = [(w[i][j], -j) for j in adj[i] if j in joined]
        assert cand, i
        _, mj = max(cand)
        assert agg[i] == agg[-mj]
    assert nagg == len(roots)
    return agg, nagg, rounds, len(roots)


def test_priority_keys_distinct():
    keys = [ML.priority_key(i, s) for s in (0, 1, 5) for i in range(20000)]
    per_salt = [set(keys[k * 20000:(k + 1) * 20000]) for k in range(3)]
    assert all(len(p) == 20000 for p in per_salt)
    assert ML.priority_key(7, 0) & 0xFFFFFFFF == 8


@pytest.mark.parametrize("m,theta", [(20, 0.08), (20, 0.0), (14, 0.25)])
def test_plain_aggregation_properties_fem(m, theta):
    a = X.gen_fem2d(m, 1.0, 3, 1e3, 0.3, 17 + m)
    A = a.to_scipy().tocsr()
    agg, nagg, rounds, nroots = check_aggregation_properties(A.indptr, A.indices, A.data, theta, 0)
    assert nagg < a.n // 2 and rounds >= 1
    sizes = np.bincount(agg)
    print(f"m {m} theta {theta}: {nagg} aggregates ({nroots} roots), sizes {sizes.min()}..{sizes.max()}, {rounds} rounds")


def brute_force(rp, ci, va, theta, salt):
    """The rounds restated with all-pairs strong distances."""
    n = len(rp) - 1
    adj, w = adjacency(rp, ci, va, theta)
    dist = [bfs_dist(adj, i) for i in range(n)]
    key = [ML.priority_key(i, salt) for i in range(n)]
    state = [0] * n
    rounds = 0
    while 0 in state:
        rounds += 1
        und = [i for i in range(n) if state[i] == 0]
        new = [i for i in und if all(key[i] > key[j] for j in und if j != i and dist[i].get(j, 3) <= 2)]
        for i in new:
            state[i] = 1
        roots = [i for i in range(n) if state[i] == 1]
        for i in und:
            if state[i] == 0 and any(dist[i].get(r, 3) <= 2 for r in roots):
                state[i] = 2
    roots = [i for i in range(n) if state[i] == 1]
    agg = [-1] * n
    for k, r in enumerate(roots):
        agg[r] = k
    for i in range(n):
        if state[i] != 1:
            for j in adj[i]:
                if state[j] == 1:
                    agg[i] = agg[j]
    first = list(agg)
    for i in range(n):
        if first[i] < 0:
            cand = [(w[i][j], -j) for j in adj[i] if first[j] >= 0]
            if cand:
                agg[i] = first[-max(cand)[1]]
    k = len(roots)
    for i in range(n):
        if agg[i] < 0:
            agg[i] = k
            k += 1
    return np.array(agg), k, rounds


@pytest.mark.parametrize("seed", range(6))
def test_plain_aggregation_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(5, 60))
    # a random symmetric matrix: sparse random couplings, some weak, a few
    # isolated rows, diagonally dominant
    B = sp.random(n, n, density=float(rng.uniform(0.03, 0.15)), random_state=seed, format="csr")
    B.data = rng.uniform(-1.0, 1.0, B.nnz) * np.where(rng.random(B.nnz) < 0.3, 0.01, 1.0)
    S = (B + B.T).tocsr()
    S.setdiag(0.0)
    S.eliminate_zeros()
    A = (S + sp.diags(np.asarray(abs(S).sum(axis=1)).ravel() + 1.0)).tocsr()
    A.sort_indices()
    for theta, salt in ((0.0, seed), (0.08, 3), (0.3, 1)):
        agg, nagg, rounds = ML.plain_aggregation(A.indptr, A.indices, A.data, theta, salt)
        want, wn, wr = brute_force(A.indptr, A.indices, A.data, theta, salt)
        np.testing.assert_array_equal(agg, want)
        assert (nagg, rounds) == (wn, wr)


def test_pieces_partition_each_aggregate():
    a = X.gen_fem2d(24, 1.0, 3, 1e3, 0.3, 9)
    A = a.to_scipy().tocsr()
    agg, nagg, _ = ML.plain_aggregation(A.indptr, A.indices, A.data, 0.0, 2)
    sl = ML.pieces(agg, nagg)
    rows = sl[sl >= 0]
    assert sorted(rows.tolist()) == list(range(a.n))       # a partition
    sizes = np.bincount(agg)
    for u, grp in enumerate(sl):
        g = grp[grp >= 0]
        assert 1 <= len(g) <= 8 and np.all(np.diff(g) > 0)  # consecutive members, ascending
        assert np.all(agg[g] == agg[g[0]])                  # within one aggregate
        assert np.all(grp[len(g):] == -1)
    # piece sizes of aggregate a: ceil(size / 8) pieces differing by at most one
    for k in range(nagg):
        ps = [int((grp >= 0).sum()) for grp in sl if grp[0] >= 0 and agg[grp[0]] == k]
        assert len(ps) == -(-sizes[k] // 8) and sum(ps) == sizes[k] and max(ps) - min(ps) <= 1
        assert ps == sorted(ps, reverse=True)
    # pieces follow aggregate order
    first = [agg[grp[0]] for grp in sl]
    assert first == sorted(first)
    # a hand case: one aggregate of 17 rows -> 6, 6, 5
    agg17 = np.zeros(17, np.int32)
    assert [int((g >= 0).sum()) for g in ML.pieces(agg17, 1)] == [6, 6, 5]


def test_galerkin_matches_scipy_product():
    a = X.gen_fem2d(30, 1.0, 3, 1e3, 0.3, 7)
    A = a.to_scipy().tocsr()
    agg, nagg, _ = ML.plain_aggregation(A.indptr, A.indices, A.data, THETA, 0)
    rp, ci, va = ML.galerkin(A.indptr, A.indices, A.data, agg, nagg)
    Ac = sp.csr_matrix((va, ci, rp), shape=(nagg, nagg)).toarray()
    P = sp.csr_matrix((np.ones(a.n), (np.arange(a.n), agg)), shape=(a.n, nagg))
    want = (P.T @ A @ P).toarray()
    assert np.all(np.diff(rp) > 0)
    for u in range(nagg):
        assert list(ci[rp[u]:rp[u + 1]]) == sorted(ci[rp[u]:rp[u + 1]])
    np.testing.assert_allclose(Ac, want, rtol=1e-13, atol=1e-13 * np.abs(want).max())


def test_galerkin_hand_case_bit_for_bit():
    # rows 0..5, aggregates {0, 3, 4} and {1, 2, 5}
    A = np.array([[4.0, -1.0, 0.0, -1.0, 0.0, 0.0],
                  [-1.0, 4.0, -1.0, 0.0, -1.0, 0.0],
                  [0.0, -1.0, 4.0, 0.0, 0.0, -1.0],
                  [-1.0, 0.0, 0.0, 4.0, -1.0, 0.0],
                  [0.0, -1.0, 0.0, -1.0, 4.0, -1.0],
                  [0.0, 0.0, -1.0, 0.0, -1.0, 4.0]]) * 0.1
    M = sp.csr_matrix(A)
    agg = np.array([0, 1, 1, 0, 0, 1], np.int32)
    rp, ci, va = ML.galerkin(M.indptr, M.indices, M.data, agg, 2)
    assert list(rp) == [0, 2, 4] and list(ci) == [0, 1, 0, 1]
    # (0,0): row 0 -> 0.4 (col 0), -0.1 (col 3); row 3 -> -0.1 (col 0), 0.4, -0.1; row 4 -> -0.1 (3), 0.4
    v00 = 0.4 + -0.1 + -0.1 + 0.4 + -0.1 + -0.1 + 0.4
    # (0,1): row 0 -> -0.1 (col 1); row 4 -> -0.1 (col 1), -0.1 (col 5)
    v01 = -0.1 + -0.1 + -0.1
    # (1,0): row 1 -> -0.1 (col 0), -0.1 (col 4); row 2 -> nothing; row 5 -> -0.1 (col 4)
    v10 = -0.1 + -0.1 + -0.1
    # (1,1): row 1 -> 0.4, -0.1 (2); row 2 -> -0.1 (1), 0.4, -0.1 (5); row 5 -> -0.1 (2), 0.4
    v11 = 0.4 + -0.1 + -0.1 + 0.4 + -0.1 + -0.1 + 0.4
    assert va.tolist() == [v00, v01, v10, v11]


def test_two_level_cycle_matches_dense_formula():
    a = X.gen_fem2d(12, 1.0, 2, 1e3, 0.3, 3)
    H = hierarchy(a, 1)
    L = H.levels[0]
    A = L.A.toarray()
    n = A.shape[0]
    M = np.zeros((n, n))
    for k, grp in enumerate(L.slots):
        g = [int(r) for r in grp if r >= 0]
        M[np.ix_(g, g)] = L.blocks[k][:len(g), :len(g)]
    V = ML.two_level_dense(A, M, L.P.toarray(), H.coarse, H.omega, H.gamma)
    rng = np.random.default_rng(1)
    for _ in range(3):
        r = rng.standard_normal(n)
        np.testing.assert_allclose(ML.vcycle(H, r), V @ r, rtol=1e-11, atol=1e-11 * np.abs(V @ r).max())


@pytest.mark.parametrize("seed_dist,omega", [(2, 0.7), (1, 0.7), (1, 0.8)])
def test_cycle_is_symmetric_positive_definite(seed_dist, omega):
    a = X.gen_fem2d(24, 1.0, 3, 1e3, 0.3, 11)
    H = hierarchy(a, 2, omega=omega, seed_dist=seed_dist)
    n = a.n
    V = np.column_stack([ML.vcycle(H, e) for e in np.eye(n)])
    assert np.abs(V - V.T).max() <= 1e-10 * np.abs(V).max()
    w = np.linalg.eigvalsh(0.5 * (V + V.T))
    assert w.min() > 0.0, w.min()


def test_pcg_with_cycle_solves_the_system():
    a = X.gen_fem2d(20, 1.0, 3, 1e3, 0.3, 5)
    H = hierarchy(a, 2)
    A = a.to_scipy().tocsr()
    x, it = ML.pcg(A, a.b, lambda r: ML.vcycle(H, r), tau=1e-28 * float(a.b @ a.b))
    want = np.linalg.solve(A.toarray(), a.b)
    assert np.linalg.norm(x - want) <= 1e-9 * np.linalg.norm(want)
    # the cycle preconditions: far fewer iterations than block Jacobi alone
    Lv = H.levels[0]
    _, it_bj = ML.pcg(A, a.b, Lv.smooth, tau=1e-12 * float(a.b @ a.b))
    _, it_ml = ML.pcg(A, a.b, lambda r: ML.vcycle(H, r), tau=1e-12 * float(a.b @ a.b))
    assert it_ml < 0.6 * it_bj, (it_ml, it_bj)

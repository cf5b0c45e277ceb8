"""Pins of the multilevel oracle (oracle/multilevel.py), on the CPU.

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

import numpy as np
import pytest
import scipy.sparse as sp

import oracle_lib as O
import xplus_amd as X

sys.path.insert(0, str(O.ROOT / "oracle"))
import blocking as BK  # noqa: E402
import multilevel as ML  # noqa: E402


def hierarchy(a, nlev, omega=0.7, gamma=1.5, seed_dist=2):
    """Levels from the rule-3 (then rule seed_dist) aggregation oracle, exact
    block inverses (FP64, no FP16 rounding), dense coarsest."""
    A = a.to_scipy().tocsr()
    rp, ci, va = A.indptr, A.indices, A.data
    levels = []
    for lev in range(nlev):
        groups = BK.parallel_blocking(list(rp), list(ci), list(va), seed_dist=2 if lev == 0 else seed_dist)
        slots = np.full((len(groups), 8), -1, np.int32)
        for k, g in enumerate(groups):
            slots[k, :len(g)] = g
        Am = sp.csr_matrix((va, ci, rp), shape=(len(rp) - 1, len(rp) - 1))
        levels.append(ML.Level(Am, slots, ML.exact_blocks(Am, slots, f32_values=False)))
        rp, ci, va = ML.galerkin(rp, ci, va, slots)
    coarse = ML.to_dense(rp, ci, va, len(rp) - 1)
    return ML.Hierarchy(levels, coarse, omega, gamma)


def test_galerkin_matches_scipy_product():
    a = X.gen_fem2d(30, 1.0, 3, 1e3, 0.3, 7)
    A = a.to_scipy().tocsr()
    groups = BK.parallel_blocking(list(A.indptr), list(A.indices), list(A.data))
    slots = np.full((len(groups), 8), -1, np.int32)
    for k, g in enumerate(groups):
        slots[k, :len(g)] = g
    rp, ci, va = ML.galerkin(A.indptr, A.indices, A.data, slots)
    Ac = sp.csr_matrix((va, ci, rp), shape=(len(groups), len(groups))).toarray()
    P = ML.Level(A, slots, np.zeros((len(groups), 8, 8))).P
    want = (P.T @ A @ P).toarray()
    assert np.all(np.diff(rp) > 0)
    for u in range(len(groups)):
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
    slots = np.array([[0, 3, 4, -1, -1, -1, -1, -1], [1, 2, 5, -1, -1, -1, -1, -1]], np.int32)
    rp, ci, va = ML.galerkin(M.indptr, M.indices, M.data, slots)
    assert list(rp) == [0, 2, 4] and list(ci) == [0, 1, 0, 1]
    # (0,0): row 0 -> 0.4 (col 0), -0.1 (col 3); row 3 -> -0.1 (col 0), 0.4, -0.1; row 4 -> -0.1 (3), 0.4
    v00 = 0.4 + -0.1 + -0.1 + 0.4 + -0.1 + -0.1 + 0.4
    # (0,1): row 0 -> -0.1 (col 1); row 4 -> -0.1 (col 1), -0.1 (col 5)
    v01 = -0.1 + -0.1 + -0.1
    v10 = -0.1 + -0.1 + -0.1
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

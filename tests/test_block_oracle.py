"""Pins the CPU oracle of the learned block preconditioner (oracle/block_mlp.py)
and checks the library's host-side block features against it (no GPU)."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle_lib as O
import xplus_amd as X

sys.path.insert(0, str(O.ROOT / "oracle"))
import block_mlp as B  # noqa: E402


def small_set(seed=0, m=18):
    a = X.gen_fem2d(m, 1.0, 3, 1e3, 0.3, seed)
    return a, B.features_from_csr(a.row_ptr, a.col_idx, a.values)


def random_params(seed):
    rng = np.random.default_rng(seed)
    p = B.unpack(np.zeros(B.N_PARAMS))
    p["W1"] = rng.standard_normal(p["W1"].shape) * 0.2
    p["b1"] = rng.standard_normal(p["b1"].shape) * 0.1 + 0.05
    p["W2"] = rng.standard_normal(p["W2"].shape) * 0.15
    p["b2"] = rng.standard_normal(p["b2"].shape) * 0.1 + 0.05
    p["W3"] = rng.standard_normal(p["W3"].shape) * 0.1
    p["b3"] = rng.standard_normal(p["b3"].shape) * 0.1
    return B.pack(p)


def test_oracle_gradient_matches_finite_differences():
    a, (feat, ahat, hmat, scale) = small_set(1)
    params = random_params(2)
    loss, g = B.loss_grad(params, feat, ahat, hmat)
    rng = np.random.default_rng(3)
    idx = rng.choice(B.N_PARAMS, 60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros(B.N_PARAMS)
        e[i] = h
        lp, _ = B.loss_grad(params + e, feat, ahat, hmat)
        lm, _ = B.loss_grad(params - e, feat, ahat, hmat)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    assert loss > 0


def test_oracle_loss_closed_form():
    """tr(M H M^T) - 2 tr(M A) + 8 equals ||M A - I||^2 + LAMBDA ||M A_off||^2 written out."""
    a, (feat, ahat, hmat, scale) = small_set(4)
    params = random_params(5)
    out = B.forward(params, feat)
    L = B.factor(out, feat)
    M = L @ np.transpose(L, (0, 2, 1)) + 1e-3 * np.eye(8)
    g = hmat - ahat @ ahat                     # = LAMBDA S_off S_off^T
    direct = [np.linalg.norm(M[k] @ ahat[k] - np.eye(8)) ** 2 + np.trace(M[k] @ g[k] @ M[k].T)
              for k in range(len(M))]
    loss, _ = B.loss_grad(params, feat, ahat, hmat)
    assert loss == pytest.approx(np.mean(direct), rel=1e-10)
    # G is PSD (a Gram matrix of the off-block couplings)
    assert min(np.linalg.eigvalsh(gk).min() for gk in g) > -1e-10


def test_library_features_match_oracle():
    a, (feat, ahat, hmat, scale) = small_set(6, m=27)   # n = 625: a partial last block
    assert a.n % 8 != 0
    bs = X.blockset(a)
    f, ah, hm, sc = bs.arrays()
    assert len(bs) == (a.n + 7) // 8
    np.testing.assert_allclose(f, feat, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(ah, ahat, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(hm, hmat, rtol=1e-5, atol=1e-6 * np.abs(hmat).max())
    np.testing.assert_allclose(sc, scale, rtol=1e-6)


def test_features_hand_block():
    """One 8x8 tridiagonal block with a single off-block neighbour: closed forms."""
    n = 9
    rp, ci, va = [0], [], []
    for r in range(n):
        for c in (r - 1, r, r + 1):
            if 0 <= c < n:
                ci.append(c)
                va.append(4.0 if c == r else -1.0)
        rp.append(len(ci))
    feat, ahat, hmat, scale = B.features_from_csr(np.array(rp), np.array(ci), np.array(va))
    np.testing.assert_allclose(scale[0], 0.5)                     # D^-1/2 = 1/2
    np.testing.assert_allclose(np.diag(ahat[0]), 1.0)               # unit diagonal
    assert ahat[0][0, 1] == -0.25
    assert feat[0, 36 + 7] == 0.25 and feat[0, 36:43].sum() == 0.0   # only row 7 couples to row 8
    g = (hmat[0] - ahat[0] @ ahat[0]) / B.LAMBDA
    assert g[7, 7] == pytest.approx(1 / 16) and abs(g).sum() == pytest.approx(1 / 16)


def test_block_pcg_diagonal_blocks_is_jacobi():
    a = X.gen_diffusion2d(16, 16, 1.0, 3)
    A = a.to_scipy()
    d = A.diagonal()
    nb = (a.n + 7) // 8
    blocks = np.zeros((nb, 8, 8))
    for i in range(a.n):
        blocks[i // 8, i % 8, i % 8] = 1.0 / d[i]
    tau = 1e-20
    x, it, tr = B.block_pcg(A, a.b, blocks, tau)
    g = O.golden(a, tau=tau)
    assert it == g["iterations"]
    np.testing.assert_allclose(tr, g["trace"][:, 4], rtol=1e-6)   # 1/d multiply vs divide, spmv order
    np.testing.assert_allclose(x, g["x"], rtol=1e-10, atol=1e-12)


def test_block_pcg_exact_inverse_on_block_diagonal_matrix():
    rng = np.random.default_rng(0)
    n = 40
    dense = np.zeros((n, n))
    for b in range(n // 8):
        q = rng.standard_normal((8, 8))
        dense[8 * b:8 * b + 8, 8 * b:8 * b + 8] = q @ q.T + 8 * np.eye(8)
    inv = np.stack([np.linalg.inv(dense[8 * b:8 * b + 8, 8 * b:8 * b + 8]) for b in range(n // 8)])
    x, it, _ = B.block_pcg(dense, np.ones(n), inv, 1e-24)
    assert it == 1
    np.testing.assert_allclose(dense @ x, np.ones(n), rtol=1e-10)


def test_oracle_l0_is_the_exact_inverse_factor():
    """L0 L0^T = Ahat^-1 (closed form), L0 lower with a positive diagonal."""
    a, (feat, ahat, hmat, scale) = small_set(6)
    L0 = B.l0_from_feat(feat)
    for k in range(len(L0)):
        assert np.allclose(np.triu(L0[k], 1), 0.0)
        assert np.all(np.diag(L0[k]) > 0)
        np.testing.assert_allclose(L0[k] @ L0[k].T @ ahat[k], np.eye(8), atol=1e-9)
    # a block that is not positive definite falls back to the identity
    bad = feat[:1].copy()
    bad[0, 0] = -1.0
    np.testing.assert_array_equal(B.l0_from_feat(bad)[0], np.eye(8))


def test_oracle_stored_blocks():
    """Diagonal blocks are stored exactly; SPD blocks within FP16 rounding of
    the unit-diagonal factor; the result is symmetric positive definite."""
    rng = np.random.default_rng(8)
    D = np.array([np.diag(rng.uniform(0.01, 100.0, 8)) for _ in range(4)])
    np.testing.assert_allclose(B.stored_blocks(D), D, rtol=1e-15)
    Q = rng.standard_normal((6, 8, 8))
    M = Q @ np.transpose(Q, (0, 2, 1)) + 0.5 * np.eye(8)
    S = B.stored_blocks(M)
    for k in range(len(M)):
        np.testing.assert_allclose(S[k], S[k].T, atol=0)
        assert np.linalg.eigvalsh(S[k]).min() > 0
        rel = np.abs(S[k] - M[k]) / np.sqrt(np.outer(np.diag(M[k]), np.diag(M[k])))
        assert rel.max() < 4e-3


def _exact_blocks(A, groups):
    out = np.zeros((len(groups), 8, 8))
    for b, g in enumerate(groups):
        k = len(g)
        out[b, :k, :k] = np.linalg.inv(A[g][:, g].toarray())
        out[b, k:, k:] = np.eye(8 - k)
    return out


def test_accurate_block_pcg_without_rounding_is_block_pcg():
    """With FP64 vectors the accurate arithmetic is block PCG with A p carried
    by its recurrence: the same iterates up to FP64 rounding on a
    well-conditioned system."""
    a = X.gen_diffusion2d(20, 18, 0.5, 31)
    A = a.to_scipy()
    groups = B.natural_groups(a.n)
    blocks = _exact_blocks(A, groups)
    tau = (1e-8 * np.linalg.norm(a.b)) ** 2
    x1, it1, _ = B.block_pcg(A, a.b, blocks, tau, groups=groups)
    x2, it2, restarts = B.block_pcg_accurate(A, a.b, blocks, tau, groups=groups, round32=False)
    assert abs(it1 - it2) <= 1 and restarts == 0
    assert np.linalg.norm(x1 - x2) <= 1e-9 * np.linalg.norm(x1)


def test_accurate_block_pcg_stops_on_the_true_residual():
    """With FP32 r and z: the returned x meets tau in the TRUE residual; the
    rounding changes the iterates against FP64 vectors; on a block-diagonal
    matrix with exact blocks one step meets 1e-6, and a 1e-12 target needs
    restarts from the true residual (each FP32 step gains ~7 digits)."""
    a = X.gen_fem2d(24, 1.0, 3, 1e3, 0.3, 32)
    A = a.to_scipy()
    groups = B.natural_groups(a.n)
    blocks = B.stored_blocks(_exact_blocks(A, groups).astype(np.float32))
    tau = (1e-6 * np.linalg.norm(a.b)) ** 2
    x, it, _ = B.block_pcg_accurate(A, a.b, blocks, tau, groups=groups)
    r = a.b - A @ x
    assert r @ r <= tau
    x64, _, _ = B.block_pcg_accurate(A, a.b, blocks, tau, groups=groups, round32=False)
    assert np.linalg.norm(x - x64) > 0.0
    import scipy.sparse as sp
    D = sp.block_diag([np.eye(8) * 2 + np.diag(np.full(7, -0.5), 1) + np.diag(np.full(7, -0.5), -1)] * 3).tocsr()
    g = B.natural_groups(24)
    bd = _exact_blocks(D, g)
    b = np.linspace(-1, 1, 24)
    xd, itd, rs = B.block_pcg_accurate(D, b, bd, (1e-6 * np.linalg.norm(b)) ** 2, groups=g)
    assert itd == 1 and rs == 0
    xd, itd, rs = B.block_pcg_accurate(D, b, bd, (1e-12 * np.linalg.norm(b)) ** 2, groups=g)
    assert rs >= 1 and itd <= 2 * (rs + 1)
    assert np.linalg.norm(D @ xd - b) <= 1e-12 * np.linalg.norm(b)


def test_bf16_rounding_known_answers():
    """Round to nearest, ties to even, at the 8-bit significand of BF16."""
    one = 1.0
    ulp = 2.0 ** -7                      # BF16 spacing at 1
    vals = np.array([one, one + ulp / 2, one + 1.5 * ulp, one + ulp / 2 + 2.0 ** -20, -3.0, 0.0])
    want = np.array([one, one, one + 2 * ulp, one + ulp, -3.0, 0.0])
    np.testing.assert_array_equal(B.bf16(vals), want)


def test_bf16_rounding_points_move_the_gradient_at_bf16_level():
    """loss_grad with the device's rounding points against the exact one:
    the same function up to BF16 rounding (relative 2^-8 per product input),
    and identical when nothing needs rounding (all-zero parameters)."""
    a, (feat, ahat, hmat, scale) = small_set(6)
    params = random_params(7)
    l0, g0 = B.loss_grad(params, feat, ahat, hmat)
    l1, g1 = B.loss_grad(params, feat, ahat, hmat, round_bf16=True)
    assert abs(l1 - l0) <= 2e-2 * abs(l0)
    assert 0 < np.linalg.norm(g1 - g0) <= 0.1 * np.linalg.norm(g0)
    z = np.zeros(B.N_PARAMS)
    lz0, gz0 = B.loss_grad(z, feat, ahat, hmat)
    lz1, gz1 = B.loss_grad(z, feat, ahat, hmat, round_bf16=True)
    assert lz0 == lz1
    np.testing.assert_array_equal(gz0, gz1)

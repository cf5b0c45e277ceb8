"""GPU tests of the learned 8x8 block-Jacobi preconditioner.

Bars:
  * MFMA forward (BF16 inputs/weights, FP32 accumulation) against the FP64
    oracle: normwise relative error <= 2e-2 per output;
  * loss and gradient (BF16 MFMA input-gradients, FP32 weight gradients):
    against oracle/block_mlp.py's loss_grad with its BF16 rounding points
    (round_bf16), normwise relative error <= 1e-3 per parameter tensor;
    against the plain FP64 loss_grad, <= 0.15 for the hidden layers and
    <= 5e-2 for W3/b3 (what BF16 rounding of the backward products' inputs
    alone moves), loss within 1e-2;
  * block-preconditioned FP32 PCG against the FP64 oracle block PCG with the
    same blocks: iterations within +-1, x within relative 1e-5, rr trace
    within 1e-4 while rr > 1e-3 rr0 (config-1 size);
  * the trained network beats plain Jacobi in iterations on held-out
    matrices (the gain is measured and printed);
  * a useless block preconditioner triggers the Jacobi fallback.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle_lib as O
import xplus_amd as X

sys.path.insert(0, str(O.ROOT / "oracle"))
import block_mlp as B  # noqa: E402
import blocking as BK  # noqa: E402

pytestmark = pytest.mark.gpu


def test_mfma_forward_matches_oracle(gpu):
    a = X.gen_fem2d(40, 1.0, 4, 1e3, 0.3, 3)
    bs = X.blockset(a)
    net = X.Mlp(seed=11)
    params = net.params().astype(np.float64)
    out = net.forward(bs)
    feat = bs.arrays()[0].astype(np.float64)
    ref = B.forward(params, feat)
    # the oracle with the device's BF16 rounding points
    ref_bf = B.forward(params, feat, round_bf16=True)
    print("forward rel err vs fp64:", np.linalg.norm(out - ref) / np.linalg.norm(ref),
          "vs bf16 model:", np.linalg.norm(out - ref_bf) / np.linalg.norm(ref_bf), flush=True)
    assert np.linalg.norm(out - ref) <= 2e-2 * np.linalg.norm(ref)
    assert np.linalg.norm(out - ref_bf) <= 1e-5 * np.linalg.norm(ref_bf) + 1e-6


def test_loss_and_gradient_match_oracle(gpu):
    a = X.gen_fem2d(34, 1.0, 4, 1e3, 0.3, 5)
    bs = X.blockset(a)
    net = X.Mlp(seed=12)
    # move away from the initialisation so every layer carries gradient
    net.train(bs, 20, batch=len(bs), lr=3e-3, seed=1)
    params = net.params().astype(np.float64)
    loss, grad = net.loss_grad(bs)
    feat, ahat, hmat, _ = (v.astype(np.float64) for v in bs.arrays())
    lref, gref = B.loss_grad(params, feat, ahat, hmat)
    _, gbf = B.loss_grad(params, feat, ahat, hmat, round_bf16=True)
    print(f"loss gpu={loss:.6e} oracle={lref:.6e}", flush=True)
    assert loss == pytest.approx(lref, rel=1e-2)
    o = 0
    for name, shp in B.SHAPES:
        k = int(np.prod(shp))
        g, gr, gb = grad[o:o + k].astype(np.float64), gref[o:o + k], gbf[o:o + k]
        rel = np.linalg.norm(g - gr) / max(np.linalg.norm(gr), 1e-30)
        relb = np.linalg.norm(g - gb) / max(np.linalg.norm(gb), 1e-30)
        print(f"grad {name}: rel err vs fp64 {rel:.3e}, vs bf16 model {relb:.3e} "
              f"(|g| = {np.linalg.norm(gr):.3e})", flush=True)
        assert relb <= 1e-3, name
        # BF16 rounding of the backward GEMM inputs moves the hidden-layer
        # gradients by several percent against exact FP64 arithmetic
        assert rel <= (5e-2 if name in ("W3", "b3") else 0.15), name
        o += k


def test_untrained_network_gives_exact_block_inverses(gpu):
    """With all parameters zero the network's correction vanishes and the
    blocks are D^-1/2 (Ahat^-1 + 1e-3 I) D^-1/2: the device's reverse
    Cholesky factor L0 against numpy's inverse."""
    a = X.gen_fem2d(30, 1.0, 3, 1e3, 0.3, 41)
    for blocking in (0, 1, 2):
        bs = X.blockset(a, blocking)
        net = X.Mlp(seed=1)
        net.set_params(np.zeros(X.Mlp.num_params(), np.float32))
        got = net.blocks(bs).astype(np.float64)
        feat, ahat, _, scale = (v.astype(np.float64) for v in bs.arrays())
        ref = (np.linalg.inv(B.ahat_from_feat(feat)) + 1e-3 * np.eye(8)) * scale[:, :, None] * scale[:, None, :]
        err = np.abs(got - ref).max() / np.abs(ref).max()
        print(f"blocking {blocking}: max rel err {err:.2e}", flush=True)
        assert err <= 1e-5


def test_training_is_deterministic_and_descends(gpu):
    sets = X.blockset([X.gen_fem2d(66, 1.0, 6, 1e3, 0.3, 20265003 + i) for i in range(2)])
    n1, n2 = X.Mlp(seed=3), X.Mlp(seed=3)
    h1 = n1.train(sets, 60, batch=1024, lr=2e-3, seed=9)
    h2 = n2.train(sets, 60, batch=1024, lr=2e-3, seed=9)
    np.testing.assert_array_equal(h1, h2)
    np.testing.assert_array_equal(n1.params(), n2.params())
    print("loss", h1[0], "->", h1[-1], "ratio", np.mean(h1[-10:]) / np.mean(h1[:5]), flush=True)
    # the network starts as the exact block inverse, so training refines a
    # good start: the loss must still fall
    assert np.mean(h1[-10:]) < 0.95 * np.mean(h1[:5])


@pytest.mark.parametrize("which,blocking", [("c1", 0), ("fem", 0), ("c1", 1), ("fem", 1), ("fem", 2)])
def test_block_pcg_matches_oracle(gpu, which, blocking):
    """Full solves on the config-1 matrix; on the 1e3-contrast FEM matrix FP32
    storage moves the iterates (as for plain Jacobi), so there the bars are
    checked at a fixed K = 20.  blocking 1 / 2: strength aggregates, rule 1 /
    2 (the oracle gets the oracle's own aggregation)."""
    a = X.gen_diffusion2d(32, 32, 1.0, 20265001) if which == "c1" else X.gen_fem2d(34, 1.0, 4, 1e3, 0.3, 7)
    tau = 1e-12 * float(a.b @ a.b)
    K = None if which == "c1" else 20
    m = X.Matrix(a, block_rows=8, keep_lo=True, blocking=blocking)
    groups = (BK.strength_blocking(a.row_ptr, a.col_idx, a.values, rule=blocking) if blocking
              else B.natural_groups(a.n))
    rows = m.block_rows()
    assert [[int(r) for r in row if r >= 0] for row in rows] == groups
    # exact block inverses (mode 1) and arbitrary SPD blocks from a network (mode 2)
    net = X.Mlp(seed=21)
    net_blocks = net.blocks(X.blockset(a, blocking)).astype(np.float64)
    A = a.to_scipy()
    exact = np.zeros((len(groups), 8, 8))
    for b, g in enumerate(groups):
        k = len(g)
        exact[b, :k, :k] = np.linalg.inv(A[g][:, g].toarray())
        exact[b, k:, k:] = np.eye(8 - k)
    for mode, blocks in ((1, exact), (2, net_blocks)):
        s = X.Pcg(m, tau=tau if K is None else 1e-300, trace_cap=5000)
        s.set_preconditioner(mode, None if mode == 1 else net_blocks.astype(np.float32))
        s.set_rhs(a.b)
        if K is None:
            s.run()
        else:
            s.run_fixed(K)
        st = s.status()
        # the oracle applies what the device stores: FP16 factors of the FP32 blocks
        x_ref, it_ref, tr_ref = B.block_pcg(A, a.b, B.stored_blocks(blocks.astype(np.float32)),
                                            tau if K is None else 1e-300, max_iters=K or 100000, groups=groups)
        print(f"{which} blocking {blocking} mode {mode}: gpu iters {st['iterations']} oracle {it_ref}", flush=True)
        assert st["status"] == (X.CONVERGED if K is None else X.MAX_ITER)
        assert abs(st["iterations"] - it_ref) <= 1
        x = s.x()
        assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) <= 1e-5
        tr = s.trace()
        k = min(len(tr_ref), len(tr))
        # FP32 iterates drift from FP64 ones once the residual has fallen a few
        # decades: a CPU simulation of this FP32 recursion with the untrained
        # network's blocks stays within 3e-6 of FP64 while rr > 1e-3 rr0 and
        # is 25% off by rr ~ 1e-5 rr0 (CG's rr is not monotone, so a small
        # phase shift is a large relative difference).  The trace is checked
        # over the first three decades; iterations and x are checked above.
        early = tr_ref[:k] > 1e-3 * tr_ref[0]
        assert early.sum() >= 5
        np.testing.assert_allclose(tr[:k, 3][early], tr_ref[:k][early], rtol=1e-4)


@pytest.mark.parametrize("case", ["c1", "fem", "ragged", "batch", "fem_agg", "batch_agg", "fem_ord", "batch_ord"])
def test_device_features_match_oracle(gpu, case):
    """Features computed on the GPU from the SELL matrix (solve order, mapped
    back to natural block order) against the oracle's features_from_csr."""
    if case == "c1":
        systems = [X.gen_diffusion2d(32, 32, 1.0, 20265001)]
    elif case == "fem":
        systems = [X.gen_fem2d(40, 1.0, 4, 1e3, 0.3, 31)]
    elif case == "ragged":                                  # n = 143: a partial last block
        systems = [X.gen_diffusion2d(13, 11, 1.0, 7)]
    elif case in ("fem_agg", "fem_ord"):
        systems = [X.gen_fem2d(40, 1.0, 4, 1e3, 0.3, 31)]
    else:                                                    # two systems of uneven size
        systems = [X.gen_diffusion2d(13, 11, 1.0, 8), X.gen_fem2d(30, 1.0, 3, 1e3, 0.3, 9)]
    blocking = 1 if case.endswith("_agg") else (2 if case.endswith("_ord") else 0)
    m = X.Matrix(systems, sigma=256, block_rows=8, keep_lo=True, blocking=blocking)
    feat, scale = m.block_features()
    ref_f, ref_s = [], []
    for a in systems:
        groups = BK.strength_blocking(a.row_ptr, a.col_idx, a.values, rule=blocking) if blocking else None
        f, _, _, sc = B.features_from_csr(a.row_ptr, a.col_idx, a.values, groups)
        ref_f.append(f)
        ref_s.append(sc)
    ref_f, ref_s = np.concatenate(ref_f), np.concatenate(ref_s)
    assert feat.shape == ref_f.shape
    # both are FP32 roundings of FP64 values (hi+lo pairs on the GPU)
    np.testing.assert_allclose(feat, ref_f, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(scale, ref_s, rtol=1e-6)


@pytest.mark.parametrize("blocking", [0, 1, 2])
def test_device_mlp_preconditioner_matches_host_blocks(gpu, blocking):
    """set_mlp_preconditioner (device features + inference + solve-order
    blocks) against mode 2 with host-built blocks of the same network."""
    train = X.gen_fem2d(66, 1.0, 6, 1e3, 0.3, 20265300)
    a = X.gen_fem2d(61, 1.0, 6, 1e3, 0.3, 20265301)          # n = 3481: a partial last block
    net = X.Mlp(seed=4)
    net.train(X.blockset(train, blocking), 100, batch=512, lr=2e-3, seed=2)
    m = X.Matrix(a, block_rows=8, keep_lo=True, blocking=blocking)
    res = []
    for how in ("device", "host"):
        s = X.Pcg(m, rel_tol=1e-6, stall_limit=0)
        if how == "device":
            s.set_mlp_preconditioner(net)
        else:
            s.set_preconditioner(2, net.blocks(X.blockset(a, blocking)))
        s.set_rhs(a.b)
        s.run()
        res.append((s.status(), s.x()))
    (s1, x1), (s2, x2) = res
    print("device", s1["iterations"], "host", s2["iterations"], flush=True)
    assert s1["status"] == s2["status"] == X.CONVERGED
    assert abs(s1["iterations"] - s2["iterations"]) <= 1
    assert np.linalg.norm(x1 - x2) <= 1e-5 * np.linalg.norm(x2)


def test_trained_network_beats_jacobi_on_held_out(gpu):
    """Train on FEM matrices of one seed split, compare on another, with
    blocks of consecutive rows and with strength aggregates."""
    train = [X.gen_fem2d(130, 1.0, 20, 1e3, 0.3, 20265100 + i) for i in range(3)]
    test = X.gen_fem2d(130, 1.0, 20, 1e3, 0.3, 20265200)
    tau = (1e-6 * np.linalg.norm(test.b)) ** 2
    results = {}
    for blocking in (0, 1):
        net = X.Mlp(seed=5)
        hist = net.train(X.blockset(train, blocking), 400, batch=4096, lr=2e-3, seed=1)
        print(f"blocking {blocking}: train loss", hist[0], "->", hist[-1], flush=True)
        m = X.Matrix(test, block_rows=8, keep_lo=True, blocking=blocking)
        for name in ("jacobi", "exact_block", "learned"):
            s = X.Pcg(m, tau=tau, stall_limit=0)
            if name == "exact_block":
                s.set_preconditioner(1)
            elif name == "learned":
                s.set_mlp_preconditioner(net)
            s.set_rhs(test.b)
            s.run()
            st = s.status()
            assert st["status"] == X.CONVERGED, (name, st)
            results[(name, blocking)] = st["iterations"]
    print("held-out iterations:", results, "gain learned vs jacobi: strips %.3f, aggregates %.3f" %
          (results[("jacobi", 0)] / results[("learned", 0)], results[("jacobi", 1)] / results[("learned", 1)]),
          flush=True)
    assert results[("learned", 0)] < results[("jacobi", 0)]
    assert results[("learned", 1)] < results[("learned", 0)]


def test_useless_blocks_fall_back_to_jacobi(gpu):
    a = X.gen_fem2d(34, 1.0, 4, 1e3, 0.3, 8)
    m = X.Matrix(a, block_rows=8)
    nb = m.num_blocks()
    rng = np.random.default_rng(0)
    q = rng.standard_normal((nb, 8, 8))
    bad = (q @ np.transpose(q, (0, 2, 1)) + 1e-3 * np.eye(8)) * 1e-2   # SPD, unrelated to A
    s = X.Pcg(m, rel_tol=1e-6, stall_limit=100)
    s.set_preconditioner(2, bad.astype(np.float32))
    s.set_rhs(a.b)
    s.run()
    st = s.status()
    print("fallback status", st, flush=True)
    assert st["status"] == X.CONVERGED
    assert st["fallback"] == 1


def test_trained_network_reproduces_exact_block_inverses(gpu):
    """What the network is for on these aggregates, measured: trained with
    the benchmark's loss (no off-block weight: its minimiser is the exact
    block inverse) on other matrices of the family, the operator it makes on
    a held-out matrix (read back from the device: S C C^T S of the stored
    FP16 factors) needs within 3% of the iterations of the exact block
    inverses' stored operator in the FP64 oracle block PCG, on the device's
    aggregates -- a count that does not move with the order of the GPU's
    sums (the accurate GPU solves of the same 140^2 system took 900 / 899
    iterations with one reduction order and 967 / 902 with another) -- and
    the GPU solve with the network's blocks takes far fewer iterations than
    Jacobi.  It does not beat exact 8x8 block inverses: the gain over Jacobi
    is the block structure."""
    train = [X.gen_fem2d(130, 1.0, 20, 1e3, 0.3, 20265700 + i) for i in range(3)]
    test = X.gen_fem2d(140, 1.0, 20, 1e3, 0.3, 20265710)
    sets = X.BlockSet(train[0], blocking=3, lam=0.0)
    for t in train[1:]:
        sets.append(X.BlockSet(t, blocking=3, lam=0.0))
    net = X.Mlp(seed=20265700)
    net.train(sets, 1000, batch=4096, lr=2e-3, seed=1)
    m = X.Matrix(test, block_rows=8, keep_lo=True, blocking=3)
    groups = [[int(r) for r in row if r >= 0] for row in m.block_rows()]
    A = test.to_scipy()
    tau = (1e-6 * np.linalg.norm(test.b)) ** 2
    its, oracle_its = {}, {}
    for name in ("jacobi", "exact", "learned"):
        s = X.Pcg(m, rel_tol=1e-6, accurate=True)
        if name == "exact":
            s.set_preconditioner(1)
        elif name == "learned":
            s.set_mlp_preconditioner(net)
        s.set_rhs(test.b)
        rel, _, it = s.run_reliable(1e-6, 0.0)
        assert rel[0] <= 1e-6
        its[name] = it
        if name != "jacobi":
            _, oracle_its[name], _ = B.block_pcg(A, test.b, s.block_operator(), tau, groups=groups)
        s.close()
    print("held-out iterations, GPU:", its, "FP64 oracle with the stored operators:", oracle_its, flush=True)
    assert abs(oracle_its["learned"] - oracle_its["exact"]) <= 0.03 * oracle_its["exact"]
    assert its["learned"] < 0.75 * its["jacobi"]

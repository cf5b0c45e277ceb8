"""The benchmark's solve path end to end against the oracle, at a size the
FP64 oracle finishes in seconds.

The path: strength-ordered 8-row aggregates (blocking 2 on the host, 3 on
the device), a block network
trained on other matrices of the family, its blocks computed on the device
(features, BF16 MFMA inference, FP16 factors), the accurate iteration (FP32
r and z, FP64 p / A p / x, FP64-accumulated SpMV over hi + lo16 values) and
residual replacement on the true residual (xp_pcg_run_reliable, delta 0).

Oracle: the operator the device stores for each block, read back in FP64
(xp_pcg_block_operator), drives
  * oracle/block_mlp.py::block_pcg, FP64 PCG to the same true tolerance:
    the GPU's x within relative 1e-5 of its x (a CPU run of this case puts
    the two 3e-7 apart), and the GPU's count at most 1.15x its count: FP32
    r and z cost iterations (CPU simulation of the accurate arithmetic on
    this matrix with exact blocks: 1,485-1,488 against FP64's 1,353, +10%);
  * block_pcg_accurate, the same arithmetic simulated on the CPU: the GPU's
    count within 3% (or 3) of the simulation's (two simulations that differ
    only in how the block product is rounded land 3 apart, 0.2%);
  * the FP64 residual of the GPU's x recomputed on the CPU <= 1e-6 ||b||.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle_lib as O
import xplus_amd as X

sys.path.insert(0, str(O.ROOT / "oracle"))
import block_mlp as B  # noqa: E402

pytestmark = pytest.mark.gpu

TRAIN_SEEDS = (20265500, 20265501)
HELD_OUT_SEED = 20265400


@pytest.fixture(scope="module")
def trained_net():
    if X.device_count() < 1:
        pytest.skip("no GPU visible")
    sets = X.blockset([X.gen_fem2d(130, 1.0, 20, 1e3, 0.3, s) for s in TRAIN_SEEDS], blocking=2)
    net = X.Mlp(seed=20265500)
    hist = net.train(sets, 300, batch=4096, lr=2e-3, seed=1)
    print(f"trained: loss {hist[0]:.4e} -> {hist[-1]:.4e}", flush=True)
    return net


def solve_headline(a, net, blocking, stall_limit=0):
    m = X.Matrix(a, block_rows=8, keep_lo=True, blocking=blocking)
    s = X.Pcg(m, rel_tol=1e-6, accurate=True, stall_limit=stall_limit)
    s.set_mlp_preconditioner(net)
    s.set_rhs(a.b, a.x0)
    rel, repl, iters = s.run_reliable(1e-6, 0.0)
    groups = [[int(r) for r in row if r >= 0] for row in m.block_rows()]
    out = {"rel": float(rel[0]), "repl": repl, "iters": iters, "x": s.x(), "status": s.status(),
           "blocks": s.block_operator(), "groups": groups}
    s.close()
    m.close()
    return out


@pytest.mark.parametrize("blocking", [2, 3])
def test_headline_path_matches_oracle(gpu, trained_net, blocking):
    """blocking 2: the host aggregation; 3: the parallel aggregation the
    device set-up computes (the benchmark's)."""
    a = X.gen_fem2d(202, 1.0, 20, 1e3, 0.3, HELD_OUT_SEED)
    A = a.to_scipy()
    g = solve_headline(a, trained_net, blocking=blocking)
    tau = (1e-6 * np.linalg.norm(a.b)) ** 2
    x_ref, it_ref, _ = B.block_pcg(A, a.b, g["blocks"], tau, groups=g["groups"])
    x_sim, it_sim, rs_sim = B.block_pcg_accurate(A, a.b, g["blocks"], tau, groups=g["groups"])
    res = np.linalg.norm(a.b - A @ g["x"]) / np.linalg.norm(a.b)
    dx = np.linalg.norm(g["x"] - x_ref) / np.linalg.norm(x_ref)
    print(f"gpu: {g['iters']} iterations, {g['repl']} replacements, rel {g['rel']:.3e} (cpu {res:.3e}); "
          f"FP64 oracle {it_ref}; accurate simulation {it_sim} ({rs_sim} restarts); |dx| {dx:.2e}", flush=True)
    assert g["status"]["fallback"] == 0
    assert g["rel"] <= 1e-6 and res <= 1e-6
    assert dx <= 1e-5
    assert g["iters"] <= 1.15 * it_ref
    assert abs(g["iters"] - it_sim) <= max(3, 0.03 * it_sim)


def test_true_residual_watch_leaves_a_good_solve_alone(gpu, trained_net):
    """The stagnation watch on the true residual (stall_limit > 0) only reads
    x: a solve that progresses keeps its blocks and its iteration count."""
    a = X.gen_fem2d(160, 1.0, 20, 1e3, 0.3, HELD_OUT_SEED + 1)
    g0 = solve_headline(a, trained_net, blocking=2, stall_limit=0)
    g1 = solve_headline(a, trained_net, blocking=2, stall_limit=256)
    print(f"iterations without watch {g0['iters']}, with {g1['iters']}", flush=True)
    assert g1["status"]["fallback"] == 0
    assert g1["iters"] == g0["iters"]
    np.testing.assert_array_equal(g1["x"], g0["x"])


@pytest.mark.parametrize("accurate", [True, False])
def test_useless_blocks_fall_back_on_the_true_residual(gpu, accurate):
    """SPD blocks unrelated to A through xp_pcg_run_reliable: the true
    residual stagnates, the solver switches to Jacobi from its current x and
    still reaches the tolerance in the true residual."""
    a = X.gen_fem2d(34, 1.0, 4, 1e3, 0.3, 8)
    m = X.Matrix(a, block_rows=8, keep_lo=True)
    rng = np.random.default_rng(0)
    q = rng.standard_normal((m.num_blocks(), 8, 8))
    bad = (q @ np.transpose(q, (0, 2, 1)) + 1e-3 * np.eye(8)) * 1e-2
    s = X.Pcg(m, rel_tol=1e-6, stall_limit=100, accurate=accurate, graph_iters=16)
    s.set_preconditioner(2, bad.astype(np.float32))
    s.set_rhs(a.b)
    rel, repl, iters = s.run_reliable(1e-6, 0.0 if accurate else 0.1)
    st = s.status()
    res = np.linalg.norm(a.b - a.to_scipy() @ s.x()) / np.linalg.norm(a.b)
    print(f"accurate={accurate}: fallback {st['fallback']}, {iters} iterations, rel {rel[0]:.3e}", flush=True)
    assert st["fallback"] == 1
    assert rel[0] <= 1e-6 and res <= 1.01e-6
    # the next run tries the blocks again (and falls back again)
    rel2, _, _ = s.run_reliable(1e-6, 0.0 if accurate else 0.1)
    assert s.status()["fallback"] == 1 and rel2[0] <= 1e-6


def test_mode2_blocks_must_have_the_matrix_block_count(gpu):
    a = X.gen_fem2d(20, 1.0, 2, 1e3, 0.3, 3)
    m = X.Matrix(a, block_rows=8)
    s = X.Pcg(m, rel_tol=1e-6)
    nb = m.num_blocks()
    with pytest.raises(X.XplusError, match="num_blocks x 64"):
        s.set_preconditioner(2, np.zeros((nb - 1, 8, 8), np.float32))
    eye = np.tile(np.eye(8, dtype=np.float32), (nb, 1, 1))
    s.set_preconditioner(2, eye)
    np.testing.assert_allclose(s.block_operator(), np.tile(np.eye(8), (nb, 1, 1)), atol=0)

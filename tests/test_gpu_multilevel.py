"""The multilevel preconditioner on the GPU (xp_pcg_set_multilevel) against
oracle/multilevel.py.

Bars (written in each test):
  * every level's aggregates equal the oracle's plain aggregation of the
    level above (oracle plain_aggregation: same strength test, keys, rounds
    and joins), and every coarse level's CSR equals the oracle's Galerkin
    product over them, bit for bit (integer work and FP64 sums in a fixed
    order); the coarse levels' smoother blocks equal the aggregation oracle
    (oracle/blocking.py parallel_blocking, distance-1 seeds for
    coarse_blocking 4) or, for coarse_blocking 1, the oracle's pieces of
    the level's own aggregates, block for block;
  * one V-cycle on the device (FP32 vectors, FP32 values, FP16 block
    factors, FP32 dense coarsest inverse) equals the FP64 oracle cycle over
    the device's hierarchy within 1e-4 relative in the 2-norm (FP32
    rounding); padding rows stay exactly zero; the cycle is symmetric to
    1e-4 relative (u^T V v against v^T V u);
  * a solve to true relative residual 1e-6 (accurate iteration, residual
    replacement) reaches it, its x lies within 1e-5 relative of the FP64
    oracle PCG's x with the same cycle, and its iteration count is at most
    1.3 x + 5 of the oracle's (FP32 r, z and the FP32 cycle cost
    iterations);
  * the graph-launched solve equals the directly launched one bit for bit;
  * to 1e-10 (exact FP64 values, replacement at delta 0.01): the true
    residual meets it and x lies within 1e-6 of the FP64 golden's, in under
    a tenth of the golden's iterations;
  * useless level-0 blocks under the true-residual watch fall back to
    Jacobi (the cycle turned off with them) and still converge;
  * misuse is refused with a message.
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

pytestmark = pytest.mark.gpu

OMEGA, GAMMA, THETA = 0.8, 1.5, 0.08


def build(a, coarse_max=64, coarse_blocking=4, use_graph=True, net=None, theta=THETA):
    m = X.Matrix(a, block_rows=8, keep_lo=True, blocking=3)
    s = X.Pcg(m, rel_tol=1e-6, accurate=True, use_graph=use_graph)
    s.set_rhs(a.b, a.x0)
    if net is None:
        s.set_preconditioner(1)
    else:
        s.set_mlp_preconditioner(net)
    s.set_multilevel(coarse_max, OMEGA, GAMMA, coarse_blocking, 0, theta)
    return m, s


def padded(groups):
    out = np.full((len(groups), 8), -1, np.int32)
    for k, g in enumerate(groups):
        out[k, :len(g)] = g
    return out


def oracle_hierarchy(a, m, s):
    """The FP64 oracle cycle over the device's hierarchy: level 0 uses the
    operator the device stores for the installed blocks; the coarse levels
    the FP16-factored exact inverses of their FP32 block values."""
    info = s.multilevel_info()
    A0 = a.to_scipy().tocsr()
    nlev = info["levels"]
    aggs = [s.multilevel_level(lev)[1] for lev in range(1, nlev)]
    levels = [ML.Level(A0, m.block_rows(), s.block_operator(), aggs[0], info["n"][1])]
    coarse = None
    for lev in range(1, nlev):
        C, _ = s.multilevel_level(lev)
        if lev < nlev - 1:
            Cs = sp.csr_matrix((C.values, C.col_idx, C.row_ptr), shape=(C.n, C.n))
            slots = s.multilevel_slots(lev)
            levels.append(ML.Level(Cs, slots, ML.stored(ML.exact_blocks(Cs, slots, f32_values=True)),
                                   aggs[lev], info["n"][lev + 1]))
        else:
            coarse = ML.to_dense(C.row_ptr, C.col_idx, C.values, C.n)
    return ML.Hierarchy(levels, coarse, OMEGA, GAMMA)


@pytest.mark.parametrize("coarse_blocking,theta", [(3, THETA), (4, THETA), (4, 0.0), (1, THETA)])
def test_hierarchy_matches_oracle(gpu, coarse_blocking, theta):
    a = X.gen_fem2d(130, 1.0, 20, 1e3, 0.3, 20265700)
    m, s = build(a, coarse_max=40, coarse_blocking=coarse_blocking, theta=theta)
    info = s.multilevel_info()
    print(f"hierarchy: {info}, rounds {s.multilevel_rounds()}", flush=True)
    assert info["levels"] >= 4 and info["n"][0] == a.n and info["n"][-1] <= 40
    prev = a
    for lev in range(1, info["levels"]):
        C, agg = s.multilevel_level(lev)
        want, nagg, rounds = ML.plain_aggregation(prev.row_ptr, prev.col_idx, prev.values, theta, lev - 1)
        assert nagg == C.n and rounds == s.multilevel_rounds()[lev - 1]
        np.testing.assert_array_equal(agg, want)
        rp, ci, va = ML.galerkin(prev.row_ptr, prev.col_idx, prev.values, agg, nagg)
        np.testing.assert_array_equal(C.row_ptr, rp)
        np.testing.assert_array_equal(C.col_idx, ci)
        np.testing.assert_array_equal(C.values, va)
        assert C.n == info["n"][lev] and C.nnz == info["nnz"][lev]
        if lev < info["levels"] - 1:
            if coarse_blocking == 1:
                # pieces of the level's own aggregates (the next coarse space)
                nxt, nn, _ = ML.plain_aggregation(C.row_ptr, C.col_idx, C.values, theta, lev)
                want_slots = ML.pieces(nxt, nn)
            else:
                want_slots = padded(BK.parallel_blocking(list(C.row_ptr), list(C.col_idx), list(C.values),
                                                         seed_dist=1 if coarse_blocking == 4 else 2))
            np.testing.assert_array_equal(s.multilevel_slots(lev), want_slots)
        prev = C
    s.close()
    m.close()


def test_vcycle_matches_oracle(gpu):
    a = X.gen_fem2d(130, 1.0, 20, 1e3, 0.3, 20265701)
    m, s = build(a, coarse_max=40)
    H = oracle_hierarchy(a, m, s)
    perm = m.permutation()
    real = perm >= 0
    rng = np.random.default_rng(5)
    for trial in range(3):
        r_nat = rng.standard_normal(a.n) if trial < 2 else a.b.copy()
        r = np.zeros(perm.size, np.float32)
        r[real] = r_nat[perm[real]]
        z = s.multilevel_apply(r)
        assert np.all(z[~real] == 0.0)
        z_nat = np.zeros(a.n)
        z_nat[perm[real]] = z[real]
        rn = np.zeros(a.n)
        rn[perm[real]] = r[real]
        want = ML.vcycle(H, rn)
        err = np.linalg.norm(z_nat - want) / np.linalg.norm(want)
        print(f"trial {trial}: relative difference {err:.3e}", flush=True)
        assert err <= 1e-4
    # symmetry of the device cycle
    u = np.zeros(perm.size, np.float32)
    v = np.zeros(perm.size, np.float32)
    u[real] = rng.standard_normal(real.sum())
    v[real] = rng.standard_normal(real.sum())
    vu, vv = s.multilevel_apply(u).astype(np.float64), s.multilevel_apply(v).astype(np.float64)
    lhs, rhs = float(u @ vv), float(v @ vu)
    print(f"u.Vv {lhs:.6e} v.Vu {rhs:.6e}", flush=True)
    assert abs(lhs - rhs) <= 1e-4 * np.linalg.norm(u) * np.linalg.norm(vv)
    s.close()
    m.close()


def test_multilevel_solve_matches_oracle(gpu):
    a = X.gen_fem2d(202, 1.0, 20, 1e3, 0.3, 20265702)
    m, s = build(a, coarse_max=512)
    rel, repl, iters = s.run_reliable(1e-6, 0.0)
    x = s.x()
    A = a.to_scipy().tocsr()
    tau = (1e-6 * np.linalg.norm(a.b)) ** 2
    H = oracle_hierarchy(a, m, s)
    x_ref, it_ref = ML.pcg(A, a.b, lambda r: ML.vcycle(H, r), tau)
    res = np.linalg.norm(a.b - A @ x) / np.linalg.norm(a.b)
    dx = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
    print(f"gpu {iters} iterations ({repl} replacements), rel {rel[0]:.3e} (cpu {res:.3e}); "
          f"FP64 oracle {it_ref}; |dx| {dx:.2e}; {s.multilevel_info()}", flush=True)
    assert s.status()["status"] == X.CONVERGED
    assert rel[0] <= 1e-6 and res <= 1.01e-6
    assert dx <= 1e-5
    assert iters <= 1.3 * it_ref + 5
    s.close()
    m.close()


def test_graph_and_direct_launches_agree(gpu):
    a = X.gen_fem2d(100, 1.0, 10, 1e3, 0.3, 20265703)
    out = []
    for use_graph in (True, False):
        m, s = build(a, coarse_max=128, use_graph=use_graph)
        rel, repl, iters = s.run_reliable(1e-6, 0.0)
        out.append((iters, s.x()))
        s.close()
        m.close()
    assert out[0][0] == out[1][0]
    np.testing.assert_array_equal(out[0][1], out[1][1])


def test_learned_blocks_with_multilevel(gpu):
    """The benchmark's combination: the network's blocks as the level-0
    smoother (an untrained network starts as the exact block inverse)."""
    a = X.gen_fem2d(160, 1.0, 20, 1e3, 0.3, 20265704)
    net = X.Mlp(seed=20265704)
    m, s = build(a, coarse_max=512, net=net)
    rel, repl, iters = s.run_reliable(1e-6, 0.0)
    A = a.to_scipy().tocsr()
    res = np.linalg.norm(a.b - A @ s.x()) / np.linalg.norm(a.b)
    m2 = X.Matrix(a, block_rows=8, keep_lo=True, blocking=3)
    s2 = X.Pcg(m2, rel_tol=1e-6, accurate=True)
    s2.set_rhs(a.b, a.x0)
    s2.set_mlp_preconditioner(net)
    _, _, it_bj = s2.run_reliable(1e-6, 0.0)
    print(f"learned blocks + multilevel: {iters} iterations, rel {rel[0]:.3e}; block Jacobi alone {it_bj}", flush=True)
    assert rel[0] <= 1e-6 and res <= 1.01e-6
    assert iters < 0.3 * it_bj
    for o in (s, s2, m, m2):
        o.close()


def test_multilevel_misuse_refused(gpu):
    a = X.gen_fem2d(40, 1.0, 4, 1e3, 0.3, 20265705)
    m = X.Matrix(a, block_rows=8, keep_lo=True, blocking=3)
    s = X.Pcg(m, rel_tol=1e-6, accurate=True)
    s.set_rhs(a.b, a.x0)
    with pytest.raises(X.XplusError, match="block preconditioner"):
        s.set_multilevel()                       # no blocks installed yet
    s.set_preconditioner(1)
    with pytest.raises(X.XplusError, match="coarse_max"):
        s.set_multilevel(coarse_max=5000)
    with pytest.raises(X.XplusError, match="omega"):
        s.set_multilevel(omega=2.5)
    s.set_multilevel(coarse_max=64)
    assert s.multilevel_info()["levels"] >= 2
    s.set_multilevel(coarse_max=0)               # removed
    assert s.multilevel_info()["levels"] == 0
    f = X.Pcg(m, rel_tol=1e-6, fused=True)
    f.set_rhs(a.b, a.x0)
    f.set_preconditioner(1)
    with pytest.raises(X.XplusError, match="accurate iteration"):
        f.set_multilevel()
    with pytest.raises(X.XplusError, match="theta"):
        s.set_multilevel(theta=1.5)
    mh = X.Matrix(a, block_rows=8, keep_lo=True, blocking=0)
    h = X.Pcg(mh, rel_tol=1e-6, accurate=True)
    h.set_rhs(a.b, a.x0)
    h.set_preconditioner(1)
    with pytest.raises(X.XplusError, match="set up on the device"):
        h.set_multilevel()
    for o in (s, f, h, m, mh):
        o.close()


def test_multilevel_to_1e10_against_fp64_golden(gpu):
    """The 1e-10 path (exact FP64 values, residual replacement at delta
    0.01, restarts from replaced residuals) with the cycle: the true relative
    residual meets 1e-10 (recomputed on the CPU within 5%: the solver's
    residual uses the FP64 values too, summed in another order), the
    solution lies within 1e-6 of the FP64 golden Jacobi-PCG's (both at true
    1e-10: the golden's own error bound at this tolerance and conditioning),
    in under a tenth of the golden's iterations."""
    a = X.gen_fem2d(202, 1.0, 20, 1e3, 0.3, 20263203)
    tol = 1e-10
    g = O.golden(a, tau=(tol * np.linalg.norm(a.b)) ** 2, trace_cap=1)
    m = X.Matrix(a, keep64=True, block_rows=8, blocking=3)
    s = X.Pcg(m, rel_tol=tol, accurate=True, stall_limit=0)
    s.set_rhs(a.b, a.x0)
    s.set_preconditioner(1)
    s.set_multilevel(512, OMEGA, GAMMA, 4, 0, THETA)
    rel, repl, iters = s.run_reliable(tol, 0.01)
    x = s.x()
    A = a.to_scipy().tocsr()
    res = np.linalg.norm(a.b - A @ x) / np.linalg.norm(a.b)
    err = np.linalg.norm(x - g["x"]) / np.linalg.norm(g["x"])
    print(f"multilevel to 1e-10: {iters} iterations (golden {g['iterations']}), {repl} replacements, "
          f"rel {rel[0]:.3e} (cpu {res:.3e}), error vs golden {err:.3e}", flush=True)
    assert s.status()["status"] == X.CONVERGED
    assert rel[0] <= tol and res <= 1.05 * tol
    assert err <= 1e-6
    assert iters < 0.1 * g["iterations"]
    s.close()
    m.close()


def test_useless_blocks_fall_back_with_the_cycle(gpu):
    """A level-0 smoother that is useless (every block scaled by 1e-6: SPD,
    but the damped smoothing and the cycle then do almost nothing) under the
    true-residual stagnation watch: the solve falls back to Jacobi, which
    turns the cycle off with the blocks, and still meets the tolerance."""
    a = X.gen_fem2d(60, 1.0, 6, 1e3, 0.3, 20265706)
    m = X.Matrix(a, block_rows=8, keep_lo=True, blocking=3)
    s = X.Pcg(m, rel_tol=1e-6, accurate=True, stall_limit=16, max_iters=20000)
    s.set_rhs(a.b, a.x0)
    nb = m.num_blocks()
    blocks = np.tile(np.eye(8, dtype=np.float32) * 1e-6, (nb, 1, 1))
    s.set_preconditioner(2, blocks)
    s.set_multilevel(64, OMEGA, GAMMA, 4, 0, THETA)
    rel, repl, iters = s.run_reliable(1e-6, 0.0)
    st = s.status()
    print(f"useless blocks with the cycle: {iters} iterations, fallback {st['fallback']}, rel {rel[0]:.3e}", flush=True)
    assert st["fallback"] == 1
    assert rel[0] <= 1e-6
    s.close()
    m.close()

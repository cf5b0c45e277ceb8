"""The matrix set-up on the device (xp_matrix_create_device, blocking 3 and 4).

Bars (all exact: the set-up is integer / byte work and fixed-order FP64):
  * the parallel strength aggregation equals oracle/blocking.py's
    parallel_blocking block for block (rows, order, padding), with
    distance-2 seeds (blocking 3) and distance-1 seeds (blocking 4), on
    sparse and on dense rows (more neighbouring blocks than the growth
    bids' LDS list holds: the fallback scan);
  * the SELL arrays built on the device equal the host builder's
    (xp_sell_build_host) for the same CSR and blocks, entry for entry:
    permutation, slice pointers and bases, FP32 values, FP32 and 16-bit
    remainders, FP64 values, 16- or 32-bit columns;
  * a solve on a device-built matrix equals the same solve on the
    host-built one bit for bit (same arrays, deterministic kernels);
  * malformed CSR input is refused with the host builder's messages.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle_lib as O
import xplus_amd as X

sys.path.insert(0, str(O.ROOT / "oracle"))
import blocking as BK  # noqa: E402

pytestmark = pytest.mark.gpu


def random_spd(n, seed, density):
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    M = sp.random(n, n, density=density, random_state=rng, format="csr")
    M = M + M.T
    M.data = -np.abs(M.data)
    d = np.asarray(abs(M).sum(axis=1)).ravel() + 1.0 + rng.random(n)
    M = (M + sp.diags(d)).tocsr()
    M.sort_indices()
    return X.Csr(M.indptr.astype(np.int32), M.indices.astype(np.int32), M.data.astype(np.float64),
                 rng.uniform(-1, 1, n), np.zeros(n))


def isolated_rows():
    """A 5-point matrix with every 7th row decoupled (diagonal only)."""
    import scipy.sparse as sp
    a = X.gen_diffusion2d(17, 9, 1.0, 5)
    A = a.to_scipy().tolil()
    for r in range(0, a.n, 7):
        for c in list(A.rows[r]):
            if c != r:
                A[r, c] = 0.0
                A[c, r] = 0.0
    A = sp.csr_matrix(A)
    A.eliminate_zeros()
    A.sort_indices()
    return X.Csr(A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data, a.b, a.x0)


CASES = {
    "fem40": lambda: X.gen_fem2d(40, 1.0, 4, 1e3, 0.3, 31),
    "fem130": lambda: X.gen_fem2d(130, 1.0, 20, 1e3, 0.3, 20265100),
    "ragged": lambda: X.gen_diffusion2d(13, 11, 1.0, 7),
    "diff3d": lambda: X.gen_diffusion3d(12, 1.5, 9),
    "random": lambda: random_spd(600, 3, 0.01),
    "isolated": isolated_rows,
    "dense": lambda: random_spd(400, 11, 0.08),
}


@pytest.mark.parametrize("blocking", [3, 4])
@pytest.mark.parametrize("case", list(CASES))
def test_device_aggregation_matches_oracle(gpu, case, blocking):
    a = CASES[case]()
    m = X.Matrix(a, block_rows=8, blocking=blocking, keep_lo=True)
    got = [[int(r) for r in row if r >= 0] for row in m.block_rows()]
    want = BK.parallel_blocking(list(a.row_ptr), list(a.col_idx), list(a.values), seed_dist=2 if blocking == 3 else 1)
    st = m.setup_stats()
    print(f"{case} blocking {blocking}: n={a.n} blocks {len(got)} oracle {len(want)} stats {st}", flush=True)
    assert st["device_built"] and st["blocks"] == len(want)
    assert got == want


def compare_sell(dev, host):
    for k in ("n_pad", "nslices", "stored", "idx16"):
        assert dev[k] == host[k], k
    for k in ("perm", "slice_ptr", "slice_base", "val", "val_lo", "val_lo16", "val64", "col16", "col32"):
        if host.get(k) is None:
            continue
        np.testing.assert_array_equal(dev[k], host[k], err_msg=k)


@pytest.mark.parametrize("case,sigma,br,blocking,keep", [
    ("fem130", 256, 1, 0, "lo"), ("fem130", 64, 1, 0, "64"), ("fem130", 4096, 1, 0, "lo"),
    ("fem130", 4096, 8, 3, "lo"), ("fem130", 256, 8, 0, "64"), ("ragged", 64, 8, 3, "lo"),
    ("random", 256, 1, 0, "lo"), ("isolated", 1024, 8, 3, "64"), ("wide", 256, 1, 0, "lo"),
])
def test_device_sell_matches_host_builder(gpu, case, sigma, br, blocking, keep):
    if case == "wide":     # spans beyond 65535 columns: 32-bit column indices
        a = random_spd(70000, 5, 4e-5)
    else:
        a = CASES[case]()
    kw = {"keep_lo": keep == "lo", "keep64": keep == "64"}
    m = X.Matrix(a, sigma=sigma, block_rows=br, blocking=blocking, device_build=True, **kw)
    dev = m.sell_arrays()
    slots = m.block_rows() if br == 8 else None
    host = X.sell_build_host(a, sigma, block_rows=br, slots=slots, **kw)
    print(f"{case} sigma={sigma} br={br}: stored {dev['stored']} idx16 {dev['idx16']}", flush=True)
    compare_sell(dev, host)
    if case == "wide":
        assert dev["idx16"] == 0


@pytest.mark.parametrize("accurate", [False, True])
def test_solve_on_device_built_matrix_is_the_host_built_solve(gpu, accurate):
    a = X.gen_fem2d(70, 1.0, 6, 1e3, 0.3, 44)
    md = X.Matrix(a, block_rows=8, blocking=0, keep_lo=True, device_build=True)
    mh = X.Matrix(a, block_rows=8, blocking=0, keep_lo=True)
    blocking = 0
    res = []
    for m in (md, mh):
        s = X.Pcg(m, rel_tol=1e-6, accurate=accurate)
        s.set_preconditioner(1)
        s.set_rhs(a.b)
        if accurate:
            rel, _, iters = s.run_reliable(1e-6, 0.0)
        else:
            s.run()
            iters = s.status()["iterations"]
        res.append((iters, s.x()))
        s.close()
    print(f"blocking {blocking} accurate {accurate}: iterations {res[0][0]} / {res[1][0]}", flush=True)
    assert res[0][0] == res[1][0]
    np.testing.assert_array_equal(res[0][1], res[1][1])


def test_device_rhs_and_solution(gpu):
    """b from device memory and x into device memory: the same solve as the
    host-array calls."""
    a = X.gen_fem2d(50, 1.0, 4, 1e3, 0.3, 45)
    d = X.DeviceCsr(a)
    m = X.Matrix.from_device(d, block_rows=8, blocking=3, keep_lo=True)
    xs = []
    for how in ("host", "device"):
        s = X.Pcg(m, rel_tol=1e-6, accurate=True)
        if how == "host":
            s.set_rhs(a.b)
            s.run_reliable(1e-6, 0.0)
            xs.append(s.x())
        else:
            s.set_rhs_device(d.b.ptr)
            s.run_reliable(1e-6, 0.0)
            out = X.DeviceArray(a.n, np.float64)
            s.x_device(out.ptr)
            xs.append(out.numpy())
            out.close()
        s.close()
    np.testing.assert_array_equal(xs[0], xs[1])
    d.close()


def test_device_build_refuses_malformed_csr(gpu):
    a = X.gen_diffusion2d(6, 5, 1.0, 1)
    bad_col = X.Csr(a.row_ptr.copy(), a.col_idx.copy(), a.values.copy())
    bad_col.col_idx[5] = a.n + 3
    with pytest.raises(X.XplusError, match="col_idx out of range"):
        X.Matrix(bad_col, device_build=True)
    bad_rp = X.Csr(a.row_ptr.copy(), a.col_idx.copy(), a.values.copy())
    bad_rp.row_ptr[4], bad_rp.row_ptr[5] = bad_rp.row_ptr[5], bad_rp.row_ptr[4]
    with pytest.raises(X.XplusError, match="nondecreasing"):
        X.Matrix(bad_rp, device_build=True)
    # a missing diagonal: the matrix builds, the solver refuses it as the
    # reference does (src/CsrDataset.hpp:107-112)
    import scipy.sparse as sp
    A = a.to_scipy().tolil()
    A[7, 7] = 0.0
    A = sp.csr_matrix(A)
    A.eliminate_zeros()
    A.sort_indices()
    nodiag = X.Csr(A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data, a.b, a.x0)
    m = X.Matrix(nodiag, device_build=True, block_rows=8, blocking=3)
    with pytest.raises(X.XplusError, match="missing diagonal entry at row 7"):
        X.Pcg(m, rel_tol=1e-6)

"""The reference's Cuper packed matrix format (CPU, no GPU).

  * the library's packer (xp_cuper_pack) equals the reference's own
    build_cuper_control_matrix (oracle/_ref, compiled from
    host/cuper_control_matrix.hpp) word for word, on stencil, FEM, random,
    multi-batch and ragged matrices;
  * the pure-Python restatement (oracle/cuper_pack.py) equals the reference's
    packer where no PE list holds two entries of one column (its stable sort
    and the reference's std::sort agree there), and the row mapping and its
    inverse are checked exhaustively;
  * format invariants on every packing; import (xp_cuper_unpack) of the
    reference's packing and of the library's gives the FP32-rounded matrix.
"""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

import oracle_lib as O
import xplus_amd as X

sys.path.insert(0, str(O.ROOT / "oracle"))
import cuper_pack as P  # noqa: E402

needs_ref = pytest.mark.skipif(O.ref is None or not hasattr(O.ref, "ref_cuper_pack_info"),
                               reason="reference build oracle/_ref not present")


def csr_np(A):
    import scipy.sparse as sp
    S = sp.csr_matrix(A)
    S.sort_indices()
    n = S.shape[0]
    return X.Csr(S.indptr, S.indices, S.data, np.ones(n), np.zeros(n))


def random_sparse(n, seed, per_row=6):
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    rows = np.repeat(np.arange(n), per_row)
    cols = rng.integers(0, n, n * per_row)
    S = sp.csr_matrix((rng.standard_normal(n * per_row), (rows, cols)), shape=(n, n))
    S.sum_duplicates()
    S = S + sp.identity(n) * 10.0
    return csr_np(S)


def fp32(a):
    return a.values.astype(np.float32).astype(np.float64)


CASES = {
    "c1": lambda: X.gen_diffusion2d(32, 32, 1.0, 20261204),
    "fem": lambda: X.gen_fem2d(120, 1.0, 6, 1e3, 0.3, 4),
    "random": lambda: random_sparse(3000, 5),
    "two_batches_ragged": lambda: X.gen_diffusion2d(97, 89, 1.0, 6),          # n = 8633
    "exact_batch": lambda: X.gen_diffusion2d(128, 64, 1.0, 7),                 # n = 8192
    "multi_batch": lambda: X.gen_fem2d(150, 1.0, 4, 1e2, 0.3, 8),            # n = 21904
    "random_wide": lambda: random_sparse(20000, 9, per_row=4),
}


@needs_ref
@pytest.mark.parametrize("case", list(CASES))
def test_library_pack_equals_reference(tmp_path, case):
    a = CASES[case]()
    a.save(tmp_path)
    ptr_r, data_r = O.ref_cuper_pack(tmp_path)
    p = X.cuper_pack(a)
    print(f"{case}: n={a.n} nnz={a.nnz} batches={p.batch_count} matrix_len={p.matrix_len} "
          f"channel_size={p.channel_size}")
    np.testing.assert_array_equal(p.ptr, ptr_r)
    assert p.data.shape == data_r.shape
    assert np.array_equal(p.data, data_r)
    # import of the reference's packing
    back = X.CuperPacked(ptr_r, data_r).unpack(a.n)
    np.testing.assert_array_equal(back.row_ptr, a.row_ptr)
    np.testing.assert_array_equal(back.col_idx, a.col_idx)
    np.testing.assert_array_equal(back.values, fp32(a))


def test_row_mapping_is_a_bijection():
    """pe_for_row and the kernel's restore_row invert each other for every
    row code of the 18-bit field (rows below 2^17 checked exhaustively)."""
    seen = set()
    for r in range(1 << 17):
        pe = P.pe_for_row(r)
        ch, lane = pe // 8, pe % 8
        code = (r // 256) * 2 + (r & 1)
        assert P.restore_row(ch, lane, code) == r
        seen.add((pe, code))
    assert len(seen) == 1 << 17
    lib = X.cuper_pack(X.Csr(np.arange(1001, dtype=np.int32), np.arange(1000, dtype=np.int32), np.ones(1000)))
    assert lib.batch_count == 1


def tie_free(n, seed):
    """Diagonal plus one entry per row at column r + 2 s, 1 <= s < 128
    random: rows dealt to one PE are r, r ^ 1 and r + 256 j, so no two of
    them share a column (an even offset below 256 never reaches a row of
    the same PE, and the two rows of a pair differ in parity)."""
    rng = np.random.default_rng(seed)
    rows, cols, vals = [], [], []
    for r in range(n):
        rows.append(r)
        cols.append(r)
        vals.append(rng.uniform(1, 2))
        c = r + 2 * int(rng.integers(1, 128))
        if c < n:
            rows.append(r)
            cols.append(c)
            vals.append(rng.uniform(-1, 1))
    import scipy.sparse as sp
    a = csr_np(sp.csr_matrix((vals, (rows, cols)), shape=(n, n)))
    keys = set()
    for r in range(n):
        for c in a.col_idx[a.row_ptr[r]:a.row_ptr[r + 1]]:
            k = (P.pe_for_row(r), int(c))
            assert k not in keys
            keys.add(k)
    return a


@needs_ref
@pytest.mark.parametrize("n,seed", [(300, 1), (2000, 2)])
def test_python_oracle_equals_reference_without_ties(tmp_path, n, seed):
    a = tie_free(n, seed)
    a.save(tmp_path)
    ptr_r, data_r = O.ref_cuper_pack(tmp_path)
    ptr, data = P.pack(a.n, a.row_ptr, a.col_idx, a.values)
    assert ptr == ptr_r.tolist()
    assert np.array_equal(np.array(data, dtype=np.uint64), data_r)


@pytest.mark.parametrize("case", ["c1", "fem", "random"])
def test_invariants_and_round_trip(case):
    a = CASES[case]()
    p = X.cuper_pack(a)
    P.check_invariants(p.ptr.tolist(), p.data, a.n)
    rp, ci, va = P.unpack(p.ptr.tolist(), p.data, a.n)
    assert rp == a.row_ptr.tolist() and ci == a.col_idx.tolist()
    np.testing.assert_array_equal(np.asarray(va), fp32(a))
    # the Python restatement's own packing obeys the rules and round-trips
    ptr, data = P.pack(a.n, a.row_ptr, a.col_idx, a.values)
    P.check_invariants(ptr, data, a.n)
    assert ptr[-1] >= max(np.bincount([P.pe_for_row(r) for r in range(a.n)],
                                      weights=np.diff(a.row_ptr)).max(), 0)
    back = X.CuperPacked(ptr, np.array(data, dtype=np.uint64)).unpack(a.n)
    np.testing.assert_array_equal(back.values, fp32(a))


def test_edge_cases():
    # empty matrix; a matrix with empty rows
    p = X.cuper_pack(X.Csr(np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0)))
    assert p.batch_count == 0 and p.matrix_len == 0 and p.channel_size == 0
    a = X.Csr(np.array([0, 1, 1, 3]), np.array([0, 0, 2]), np.array([1.0, 2.0, 3.0]))
    back = X.cuper_pack(a).unpack(3)
    assert back.row_ptr.tolist() == [0, 1, 1, 3] and back.col_idx.tolist() == [0, 0, 2]


def test_import_errors():
    a = X.gen_diffusion2d(40, 40, 1.0, 1)
    p = X.cuper_pack(a)
    with pytest.raises(X.XplusError, match="outside"):
        p.unpack(a.n - 100)
    bad = X.CuperPacked(p.ptr[::-1].copy(), p.data)
    with pytest.raises(X.XplusError, match="sp_element_list_ptr"):
        bad.unpack(a.n)
    short = X.CuperPacked(p.ptr, p.data[:, :8])
    with pytest.raises(X.XplusError, match="matrix_data does not match"):
        short.unpack(a.n)
    # more rows than the 18-bit row code can address
    h = C.c_void_p()
    rp = np.zeros(2, np.int32)
    rc = X._cu_pack((1 << 24) + 1, rp, rp, np.zeros(1), C.byref(h))
    assert rc != 0 and "too large" in X.last_error()


def test_cli_rejects_bad_arguments(tmp_path):
    """xplus_cuper's argument errors (host/cuper_tapa_pcg_main.cpp:68-125):
    usage on stderr, exit code 1, before any device work."""
    import subprocess
    for args in (["--bogus"], ["--diff-tol", "0"], ["--spmv-repeats", "0"], ["--tau", "-1"], ["--max-iters"]):
        out = subprocess.run([str(X.CUPER_CLI_PATH)] + args, capture_output=True, text=True, timeout=60)
        assert out.returncode == 1 and "Usage:" in out.stderr, (args, out.stderr)
    out = subprocess.run([str(X.CUPER_CLI_PATH), str(tmp_path / "missing")], capture_output=True, text=True,
                         timeout=60)
    assert out.returncode == 1 and "failed to open" in out.stderr


def test_import_is_independent_of_slot_order():
    """Packers of this format differ in where equal columns land and in the
    slot schedule (the DLC packer, DLC/Cuper/include/Cuper_common.h:535-774,
    orders entries by its slice traversal): import must not depend on it.
    Every (channel, lane, batch) segment shuffled, empty slots included."""
    a = X.gen_fem2d(120, 1.0, 6, 1e3, 0.3, 4)
    p = X.cuper_pack(a)
    rng = np.random.default_rng(3)
    data = p.data.copy()
    for ch in range(16):
        for lane in range(8):
            for b in range(p.batch_count):
                idx = np.arange(p.ptr[b], p.ptr[b + 1]) * 8 + lane
                data[ch, idx] = data[ch, rng.permutation(idx)]
    back = X.CuperPacked(p.ptr, data).unpack(a.n)
    np.testing.assert_array_equal(back.row_ptr, a.row_ptr)
    np.testing.assert_array_equal(back.col_idx, a.col_idx)
    np.testing.assert_array_equal(back.values, fp32(a))

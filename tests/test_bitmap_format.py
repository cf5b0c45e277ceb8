"""The reference's 4x4 bitmap-block matrix format (CPU, no GPU).

  * bit layout: the example of include/cg_common.hpp:37-46 and a hand-built
    block (bit 4 r + c, values compact in bit order, mask in indices[2..3]);
  * the oracle's pack (oracle/bitmap_blocks.py) against the dense matrix,
    and its SpMV against the dense product with integer values (exact), in
    the plain and windowed layouts, and against the reference's own kernel
    (oracle/_ref/libref_blocked.so, compiled from
    archived/kernels/spmv_blocked_kernel.cpp) bit for bit;
  * the library's packer (xp_bitmap_pack) equals the oracle's byte for byte;
    its unpack (xp_bitmap_unpack) equals the oracle's and inverts the pack;
  * the in-order CSR SpMV that checks full-size matrices equals the oracle's
    SpMV on the unpacked layout;
  * malformed layouts are rejected with the documented messages.
"""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

import oracle_lib as O
import xplus_amd as X

sys.path.insert(0, str(O.ROOT / "oracle"))
import bitmap_blocks as B  # noqa: E402

_ref_blocked = C.CDLL(str(O.REF_BLOCKED_SO)) if O.REF_BLOCKED_SO.exists() else None
needs_ref = pytest.mark.skipif(_ref_blocked is None, reason="reference build oracle/_ref/libref_blocked.so not present")

WINDOWS = [0, 4, 6, 8, 512]


def csr_of(n, rows, cols, vals):
    """CSR in the given entry order (duplicates and zeros kept), b = 1."""
    order = np.argsort(rows, kind="stable")
    rows, cols, vals = np.asarray(rows)[order], np.asarray(cols)[order], np.asarray(vals, float)[order]
    rp = np.zeros(n + 1, np.int32)
    np.add.at(rp, rows + 1, 1)
    return X.Csr(np.cumsum(rp).astype(np.int32), cols.astype(np.int32), vals, np.ones(n), np.zeros(n))


def random_csr(n, seed, per_row=5, integer=False, dup=True):
    rng = np.random.default_rng(seed)
    m = n * per_row
    rows = rng.integers(0, n, m)
    cols = rng.integers(0, n, m)
    vals = rng.integers(-4, 5, m).astype(float) if integer else rng.standard_normal(m)
    if dup and m > 2:
        rows = np.r_[rows, rows[:3]]                 # duplicates of existing entries
        cols = np.r_[cols, cols[:3]]
        vals = np.r_[vals, -vals[:1], vals[1:3]]     # one of them cancels to 0.0
    return csr_of(n, rows, cols, vals)


def dense(a):
    A = np.zeros((a.n, a.n))
    for i in range(a.n):
        for o in range(a.row_ptr[i], a.row_ptr[i + 1]):
            A[i, a.col_idx[o]] += a.values[o]
    return A


def oracle_pack(a, w):
    return B.pack(a.n, a.row_ptr, a.col_idx, a.values, w)


def test_mask_layout_of_the_reference_example():
    # include/cg_common.hpp:37-46: bits 0 and 5 set -> values[0] is (0, 0),
    # values[1] is (1, 1)
    a = csr_of(4, [0, 1], [0, 1], [3.0, 7.0])
    rp, ci, blocks = oracle_pack(a, 0)
    assert list(rp) == [0, 1] and list(ci) == [0]
    assert B.mask_of(blocks[0]) == (1 << 0) | (1 << 5)
    assert list(blocks[0]["values"][:2]) == [3.0, 7.0]
    lib = X.bitmap_pack(a, 0)
    assert lib.blocks.tobytes() == blocks.tobytes()


def test_hand_built_block():
    # (0,1)=2, (0,3)=3, (2,0)=5, (3,3)=7 of a 6 x 6 matrix; (5,4)=1 in block
    # (1,1); bit 4 r + c, low byte in indices[2]
    a = csr_of(6, [0, 0, 2, 3, 5], [1, 3, 0, 3, 4], [2.0, 3.0, 5.0, 7.0, 1.0])
    lib = X.bitmap_pack(a, 0)
    assert list(lib.row_ptr) == [0, 1, 2] and list(lib.col_idx) == [0, 1]
    b0, b1 = lib.blocks[0], lib.blocks[1]
    assert b0["indices"][2] == 0x0A and b0["indices"][3] == 0x81
    assert list(b0["values"]) == [2.0, 3.0, 5.0, 7.0] + [0.0] * 12
    assert b1["indices"][2] == 0x10 and b1["indices"][3] == 0x00     # (1, 0) of block (1,1): bit 4
    assert [int(v) for i, v in enumerate(b1["indices"]) if i not in (2, 3)] == [0] * 14
    assert b1["values"][0] == 1.0


@pytest.mark.parametrize("n", [1, 3, 5, 13, 64, 101])
@pytest.mark.parametrize("w", WINDOWS)
def test_oracle_pack_reproduces_dense(n, w):
    a = random_csr(n, 100 + n, integer=True)
    rp, ci, blocks = oracle_pack(a, w)
    nbr = (n + 3) // 4
    nw = B.num_windows(n, w)
    assert len(rp) == nw * (nbr + 1)
    A = np.zeros((4 * nbr, 4 * (nbr + 1)))
    for wi in range(nw):
        base = wi * (nbr + 1)
        if wi > 0:
            assert rp[base] == rp[base - 1]                       # windows follow each other
        lo = (wi * w) // 4 if w else 0
        hi = -(-min(wi * w + w, n) // 4) if w else nbr
        for br in range(nbr):
            assert rp[base + br] <= rp[base + br + 1]
            cols = ci[rp[base + br]:rp[base + br + 1]]
            assert list(cols) == sorted(set(cols)) and all(lo <= c < hi for c in cols)
            for k in range(rp[base + br], rp[base + br + 1]):
                mask = B.mask_of(blocks[k])
                assert mask != 0
                vals = blocks[k]["values"]
                assert np.all(vals[bin(mask).count("1"):] == 0.0)
                j = 0
                for pos in range(16):
                    if mask >> pos & 1:
                        assert vals[j] != 0.0
                        r, c = divmod(pos, 4)
                        # a block column shared by two windows (w % 4 != 0)
                        # is listed in both with the same payload
                        A[4 * br + r, 4 * ci[k] + c] = vals[j]
                        j += 1
    np.testing.assert_array_equal(A[:n, :n], dense(a))
    assert not A[n:].any() and not A[:, n:].any()


@pytest.mark.parametrize("n", [1, 2, 7, 33, 130])
@pytest.mark.parametrize("w", WINDOWS)
def test_oracle_spmv_is_exact_on_integers(n, w):
    a = random_csr(n, 7 * n + w, integer=True)
    x = np.random.default_rng(n).integers(-9, 10, n).astype(float)
    rp, ci, blocks = oracle_pack(a, w)
    y = B.spmv(n, w, rp, ci, blocks, x)
    np.testing.assert_array_equal(y, dense(a) @ x)


@needs_ref
@pytest.mark.parametrize("n,seed", [(1, 0), (4, 1), (37, 2), (256, 3), (1021, 4), (1024, 5)])
def test_oracle_spmv_equals_reference_kernel(n, seed):
    a = random_csr(n, seed)
    x = np.random.default_rng(seed + 50).standard_normal(n)
    lib = X.bitmap_pack(a, 0)
    y_ref = np.full(n, np.nan)
    blocks = lib.blocks if lib.nblocks else np.zeros(1, X.SPMV_BLOCK)
    _ref_blocked.spmv_blocked_kernel(lib.row_ptr.ctypes.data_as(C.c_void_p),
                                     (lib.col_idx if lib.nblocks else np.zeros(1, np.int32)).ctypes.data_as(C.c_void_p),
                                     blocks.ctypes.data_as(C.c_void_p), x.ctypes.data_as(C.c_void_p),
                                     y_ref.ctypes.data_as(C.c_void_p), C.c_int(n))
    y = B.spmv(n, 0, lib.row_ptr, lib.col_idx, lib.blocks, x)
    assert y.tobytes() == y_ref.tobytes()


@pytest.mark.parametrize("n", [1, 6, 29, 200])
@pytest.mark.parametrize("w", WINDOWS)
def test_library_pack_equals_oracle(n, w):
    a = random_csr(n, 3 * n + w)
    rp, ci, blocks = oracle_pack(a, w)
    lib = X.bitmap_pack(a, w)
    assert lib.num_windows == B.num_windows(n, w)
    np.testing.assert_array_equal(lib.row_ptr, rp)
    np.testing.assert_array_equal(lib.col_idx, ci)
    assert lib.blocks.tobytes() == blocks.tobytes()


@pytest.mark.parametrize("w", WINDOWS)
def test_windowed_spmv_equals_plain(w):
    # finite values: the windowed order reads the plain order's products,
    # with extra +-0.0 terms for the columns outside a shared block's window
    a = random_csr(150, 11)
    x = np.random.default_rng(12).standard_normal(150)
    y0 = B.spmv(150, 0, *oracle_pack(a, 0), x)
    yw = B.spmv(150, w, *oracle_pack(a, w), x)
    assert y0.tobytes() == yw.tobytes()


def canonical_csr(n, seed):
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    S = sp.random(n, n, density=min(1.0, 6.0 / n), random_state=seed, format="csr") + sp.identity(n)
    S.sum_duplicates()
    S.sort_indices()
    S.data = rng.standard_normal(S.nnz)
    return X.Csr(S.indptr.astype(np.int32), S.indices.astype(np.int32), S.data, np.ones(n), np.zeros(n))


@pytest.mark.parametrize("n", [1, 9, 77, 300])
@pytest.mark.parametrize("w", WINDOWS)
def test_unpack_inverts_pack(n, w):
    a = canonical_csr(n, n + w)
    lib = X.bitmap_pack(a, w)
    u = lib.unpack()
    np.testing.assert_array_equal(u.row_ptr, a.row_ptr)
    np.testing.assert_array_equal(u.col_idx, a.col_idx)
    assert u.values.tobytes() == a.values.tobytes()
    orp, oci, ova = B.unpack(n, w, lib.row_ptr, lib.col_idx, lib.blocks)
    np.testing.assert_array_equal(orp, u.row_ptr)
    np.testing.assert_array_equal(oci, u.col_idx)
    assert ova.tobytes() == u.values.tobytes()


@pytest.mark.parametrize("w", WINDOWS)
def test_in_order_csr_spmv_equals_oracle_spmv(w):
    a = random_csr(120, 21)
    x = np.random.default_rng(22).standard_normal(120)
    lib = X.bitmap_pack(a, w)
    u = lib.unpack()
    y = B.csr_spmv_in_order(u.row_ptr, u.col_idx, u.values, x)
    assert y.tobytes() == B.spmv(120, w, lib.row_ptr, lib.col_idx, lib.blocks, x).tobytes()


def test_unpack_of_a_foreign_layout_keeps_its_order_and_drops_outside_entries():
    # n = 5: block columns 0 and 1 (columns 4..7; 5..7 lie outside the
    # matrix); block row 0 lists block 1 before block 0; a stored 0.0 is kept
    blocks = np.zeros(2, X.SPMV_BLOCK)
    blocks[0]["values"][:3] = [1.5, 0.0, 2.5]                 # (0,0) (0,1) (1,3): cols 4, 5, 7
    m0 = 1 << 0 | 1 << 1 | 1 << 7
    blocks[1]["values"][:1] = [4.0]                           # (0,2): col 2
    m1 = 1 << 2
    for b, m in ((blocks[0], m0), (blocks[1], m1)):
        b["indices"][2], b["indices"][3] = m & 0xFF, m >> 8
    bm = X.BitmapBlocks(5, 0, [0, 2, 2], [1, 0], blocks)
    u = bm.unpack()
    assert list(u.row_ptr) == [0, 2, 2, 2, 2, 2]
    assert list(u.col_idx) == [4, 2] and list(u.values) == [1.5, 4.0]
    x = np.arange(1.0, 6.0)
    np.testing.assert_array_equal(B.spmv(5, 0, bm.row_ptr, bm.col_idx, bm.blocks, x), [1.5 * 5 + 4.0 * 3, 0, 0, 0, 0])
    orp, oci, ova = B.unpack(5, 0, bm.row_ptr, bm.col_idx, bm.blocks)
    assert list(orp) == list(u.row_ptr) and list(oci) == [4, 2] and list(ova) == [1.5, 4.0]


@pytest.mark.parametrize("rp,ci,msg", [
    ([0, 2, 1], [0, 0], "bitmap row_ptr must be nondecreasing"),
    ([0, 1, 3], [0, 0], "bitmap row_ptr out of range"),
    ([0, 1, 2], [0, -1], "bitmap col_idx must be non-negative"),
])
def test_unpack_rejects_malformed_layouts(rp, ci, msg):
    blocks = np.zeros(2, X.SPMV_BLOCK)
    bm = X.BitmapBlocks(8, 0, rp, ci, blocks)
    with pytest.raises(RuntimeError, match=msg):
        bm.unpack()


def test_unpack_rejects_windows_that_do_not_follow_each_other():
    # n = 8, window 4: two windows of 3 pointers; window 1 starts at 0 again
    blocks = np.zeros(2, X.SPMV_BLOCK)
    bm = X.BitmapBlocks(8, 4, [0, 1, 1, 0, 1, 2], [0, 1], blocks)
    with pytest.raises(RuntimeError, match="bitmap row_ptr windows must be contiguous"):
        bm.unpack()


def test_pack_rejects_a_malformed_csr():
    a = X.Csr(np.array([0, 1], np.int32), np.array([3], np.int32), np.array([1.0]), np.ones(1), np.zeros(1))
    with pytest.raises(RuntimeError, match="col_idx out of range"):
        X.bitmap_pack(a, 0)

"""The 4x4 bitmap-block format on the GPU: the direct SpMV
(xp_bitmap_spmv_f64) and the import into a device matrix
(xp_matrix_create_bitmap).

Bars (all exact):
  * the direct SpMV equals the oracle's SpMV (oracle/bitmap_blocks.py,
    pinned against the reference's own kernel) bit for bit, in the plain
    and windowed layouts, on ragged sizes, duplicate and cancelling
    entries, empty block rows and a foreign layout whose blocks are out of
    order and reach past the matrix;
  * at the thermal-FEM 1100 x 1100 size (n = 1,205,604): bit for bit the
    in-order CSR SpMV of the unpacked layout (equal to the oracle's SpMV for
    finite values, checked on the CPU);
  * the imported matrix holds the unpacked CSR entry for entry, its SELL
    arrays equal those of the same CSR set up directly, and a solve on it
    equals that solve bit for bit;
  * malformed layouts are refused by the device import with the host
    unpack's messages.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle_lib as O
import xplus_amd as X

sys.path.insert(0, str(O.ROOT / "oracle"))
import bitmap_blocks as B  # noqa: E402
from test_bitmap_format import random_csr  # noqa: E402

pytestmark = pytest.mark.gpu


def gpu_spmv(bm, x):
    d = bm.to_device()
    try:
        return d.spmv_numpy(x)
    finally:
        d.close()


CASES = {
    "random1": lambda: random_csr(1, 1),
    "random7": lambda: random_csr(7, 2),
    "random130": lambda: random_csr(130, 3),
    "random1021": lambda: random_csr(1021, 4),
    "fem30": lambda: X.gen_fem2d(30, 1.0, 4, 1e3, 0.3, 5),
    "ragged": lambda: X.gen_diffusion2d(13, 11, 1.0, 7),
}


@pytest.mark.parametrize("w", [0, 6, 512])
@pytest.mark.parametrize("case", list(CASES))
def test_direct_spmv_equals_oracle(gpu, case, w):
    a = CASES[case]()
    bm = X.bitmap_pack(a, w)
    x = np.random.default_rng(a.n + w).standard_normal(a.n)
    y = gpu_spmv(bm, x)
    want = B.spmv(a.n, w, bm.row_ptr, bm.col_idx, bm.blocks, x)
    print(f"{case} w={w}: n={a.n} blocks {bm.nblocks}", flush=True)
    assert y.tobytes() == want.tobytes()


def test_direct_spmv_on_a_foreign_layout(gpu):
    # blocks out of column order, an empty block row, entries past n = 10
    # (they multiply 0.0) and a stored 0.0
    rng = np.random.default_rng(9)
    blocks = np.zeros(4, X.SPMV_BLOCK)
    masks = [0xFFFF, 0x8421, 0x00F0, 0x1001]
    for b, m in zip(blocks, masks):
        k = bin(m).count("1")
        b["values"][:k] = rng.standard_normal(k)
        b["indices"][2], b["indices"][3] = m & 0xFF, m >> 8
    blocks[1]["values"][1] = 0.0
    bm = X.BitmapBlocks(10, 0, [0, 2, 2, 4], [2, 0, 1, 2], blocks)
    x = rng.standard_normal(10)
    y = gpu_spmv(bm, x)
    want = B.spmv(10, 0, bm.row_ptr, bm.col_idx, bm.blocks, x)
    assert y.tobytes() == want.tobytes()


def test_direct_spmv_at_fem1100_size(gpu):
    a = X.gen_fem2d(1100, 1.0, 20, 1e3, 0.3, 20263203)
    bm = X.bitmap_pack(a, 0)
    u = bm.unpack()
    x = np.random.default_rng(1).standard_normal(a.n)
    y = gpu_spmv(bm, x)
    want = B.csr_spmv_in_order(u.row_ptr, u.col_idx, u.values, x)
    print(f"fem1100: n={a.n} nnz={a.nnz} blocks {bm.nblocks} ({a.nnz / bm.nblocks:.2f} entries per block)",
          flush=True)
    assert u.nnz == a.nnz
    assert y.tobytes() == want.tobytes()


@pytest.mark.parametrize("w,br,blocking", [(0, 1, 0), (512, 8, 3), (6, 8, 0)])
def test_import_equals_direct_set_up(gpu, w, br, blocking):
    a = X.gen_fem2d(90, 1.0, 8, 1e3, 0.3, 61)
    bm = X.bitmap_pack(a, w)
    d = bm.to_device()
    kw = dict(block_rows=br, blocking=blocking, keep_lo=True)
    m_bm = X.Matrix.from_bitmap(d, **kw)
    rp, ci, va = m_bm.csr()
    u = bm.unpack()
    np.testing.assert_array_equal(rp, u.row_ptr)
    np.testing.assert_array_equal(ci, u.col_idx)
    assert va.tobytes() == u.values.tobytes()
    dc = X.DeviceCsr(u)
    m_csr = X.Matrix.from_device(dc, **kw)
    s_bm, s_csr = m_bm.sell_arrays(), m_csr.sell_arrays()
    for k in ("n_pad", "nslices", "stored", "idx16"):
        assert s_bm[k] == s_csr[k], k
    for k in ("perm", "slice_ptr", "slice_base", "val", "val_lo", "col16", "col32"):
        if s_csr[k] is not None:
            np.testing.assert_array_equal(s_bm[k], s_csr[k], err_msg=k)
    xs = []
    for m in (m_bm, m_csr):
        s = X.Pcg(m, rel_tol=1e-6, accurate=True)
        if br == 8:
            s.set_preconditioner(1)
        s.set_rhs(a.b)
        s.run_reliable(1e-6, 0.0)
        xs.append(s.x())
        s.close()
    assert xs[0].tobytes() == xs[1].tobytes()
    m_bm.close()
    m_csr.close()
    d.close()
    dc.close()


@pytest.mark.parametrize("rp,ci,msg", [
    ([0, 2, 1], [0, 0], "bitmap row_ptr must be nondecreasing"),
    ([0, 1, 3], [0, 0], "bitmap row_ptr out of range"),
    ([0, 1, 2], [0, -1], "bitmap col_idx must be non-negative"),
])
def test_import_refuses_malformed_layouts(gpu, rp, ci, msg):
    blocks = np.zeros(2, X.SPMV_BLOCK)
    for b in blocks:
        b["values"][0] = 1.0
        b["indices"][2] = 1
    d = X.BitmapBlocks(8, 0, rp, ci, blocks).to_device()
    with pytest.raises(X.XplusError, match=msg):
        X.Matrix.from_bitmap(d)
    d.close()


def test_import_refuses_windows_that_do_not_follow_each_other(gpu):
    blocks = np.zeros(2, X.SPMV_BLOCK)
    for b in blocks:
        b["values"][0] = 1.0
        b["indices"][2] = 1
    d = X.BitmapBlocks(8, 4, [0, 1, 1, 0, 1, 2], [0, 1], blocks).to_device()
    with pytest.raises(X.XplusError, match="bitmap row_ptr windows must be contiguous"):
        X.Matrix.from_bitmap(d)
    d.close()

"""CPU oracle of the reference's 4x4 bitmap-block matrix format.

TEST INFRASTRUCTURE ONLY (imported by tests/, never by the product).
Restates, in pure Python (small cases):

  pack     build_windowed_blocked_spmv_matrix
           (archived/host/xrt_host.cpp:229-349): blocks of 4 rows x 4
           columns; per block row a map from block column to a dense 4x4
           accumulator into which the CSR entries of the block row's rows
           are added in CSR order (:262-285); per block the positions whose
           sum is != 0.0 set bit 4 r + c of the mask and append their value
           (:287-303); empty blocks dropped; mask in indices[2] (low byte)
           and indices[3] (:305-309); window-major arrangement (:323-346):
           for every window w, every block row, the blocks with
           floor(wW / 4) <= bc < ceil(min(wW + W, n) / 4), then the
           window's last pointer.  window_size 0 = the plain block-row CSR
           that the standalone kernel reads (one window, every column).
  spmv     spmv_blocked_bitmap_body (archived/kernels/spmv_blocked_kernel.cpp:
           19-89) for the plain layout, accumulate_window_tile + write_y_tile
           (archived/kernels/pcg_control_kernel.cpp:85-245) for the windowed
           one: per row an accumulator from 0.0, windows in order, blocks in
           stored order, bits ascending, acc += value * x where x is 0.0 for
           a column >= n or outside the window.
  unpack   the library's import: the entries spmv reads, in its order,
           without those outside the matrix or their window.
  csr_spmv_in_order
           y[i] = sequential sum over row i's CSR entries in their order,
           each product rounded before it is added (numpy, vectorised over
           rows): equal to spmv on unpack(layout) for finite values, and
           fast enough to check full-size matrices.

Pins (tests/test_bitmap_format.py): the bit layout against the example in
include/cg_common.hpp:37-46 and a hand-built block; pack's blocks against
the dense matrix (integer values: exact); spmv against the dense product
(integer values and x: exact) in both layouts; spmv against the
reference's own kernel (oracle/_ref/libref_blocked.so, compiled from
archived/kernels/spmv_blocked_kernel.cpp) bit for bit.  The reference's
packer itself is not buildable here (its file needs the XRT runtime
headers), so pack is pinned by those checks and by the round trip
unpack(pack(A)) = A.
"""
from __future__ import annotations

import numpy as np

SIDE = 4
ENTRIES = 16
WINDOW = 512      # the reference's kSpmvWindowSize (include/cg_common.hpp:20)
BLOCK_DTYPE = np.dtype([("values", "<f8", (16,)), ("indices", "u1", (16,))])


def num_windows(n, window_size):
    return -(-n // window_size) if window_size > 0 else 1


def mask_of(block):
    return int(block["indices"][2]) | int(block["indices"][3]) << 8


def pack(n, row_ptr, col_idx, values, window_size=WINDOW):
    """(row_ptr, col_idx, blocks) of the layout; blocks a BLOCK_DTYPE array."""
    nbr = (n + SIDE - 1) // SIDE
    rows = []                                    # per block row: [(bc, block)]
    for br in range(nbr):
        acc = {}                                 # bc -> 16 sums
        for r in range(SIDE):
            row = br * SIDE + r
            if row >= n:
                break
            for o in range(row_ptr[row], row_ptr[row + 1]):
                c = int(col_idx[o])
                dense = acc.setdefault(c // SIDE, [0.0] * ENTRIES)
                dense[r * SIDE + c % SIDE] += float(values[o])
        out = []
        for bc in sorted(acc):
            b = np.zeros((), BLOCK_DTYPE)
            mask, k = 0, 0
            for pos in range(ENTRIES):
                if acc[bc][pos] != 0.0:
                    mask |= 1 << pos
                    b["values"][k] = acc[bc][pos]
                    k += 1
            if mask == 0:
                continue
            b["indices"][2] = mask & 0xFF
            b["indices"][3] = mask >> 8
            out.append((bc, b))
        rows.append(out)

    rp, ci, blocks = [], [], []
    if window_size == 0:
        rp.append(0)
        for out in rows:
            for bc, b in out:
                ci.append(bc)
                blocks.append(b)
            rp.append(len(blocks))
    else:
        for w in range(num_windows(n, window_size)):
            wb = w * window_size
            we = min(wb + window_size, n)
            lo, hi = wb // SIDE, (we + SIDE - 1) // SIDE
            for out in rows:
                rp.append(len(blocks))
                for bc, b in out:
                    if lo <= bc < hi:
                        ci.append(bc)
                        blocks.append(b)
            rp.append(len(blocks))
    arr = np.zeros(len(blocks), BLOCK_DTYPE)
    for k, b in enumerate(blocks):
        arr[k] = b
    return np.array(rp, np.int32), np.array(ci, np.int32), arr


def _row_entries(n, window_size, row_ptr, col_idx, blocks, row):
    """(column or -1, value) of every entry y[row] reads, in order."""
    nbr = (n + SIDE - 1) // SIDE
    br, r = divmod(row, SIDE)
    for w in range(num_windows(n, window_size) if n > 0 else 0):
        wb = w * window_size if window_size > 0 else 0
        we = min(wb + window_size, n) if window_size > 0 else n
        base = w * (nbr + 1)
        for k in range(row_ptr[base + br], row_ptr[base + br + 1]):
            mask = mask_of(blocks[k])
            vi = 0
            for pos in range(ENTRIES):
                if (mask >> pos) & 1:
                    if pos // SIDE == r:
                        gc = int(col_idx[k]) * SIDE + pos % SIDE
                        yield (gc if wb <= gc < we else -1), float(blocks[k]["values"][vi])
                    vi += 1


def spmv(n, window_size, row_ptr, col_idx, blocks, x):
    y = np.zeros(n)
    for row in range(n):
        acc = 0.0
        for col, v in _row_entries(n, window_size, row_ptr, col_idx, blocks, row):
            acc = acc + v * (float(x[col]) if col >= 0 else 0.0)
        y[row] = acc
    return y


def unpack(n, window_size, row_ptr, col_idx, blocks):
    rp, ci, va = [0], [], []
    for row in range(n):
        for col, v in _row_entries(n, window_size, row_ptr, col_idx, blocks, row):
            if col >= 0:
                ci.append(col)
                va.append(v)
        rp.append(len(ci))
    return np.array(rp, np.int32), np.array(ci, np.int32), np.array(va, np.float64)


def csr_spmv_in_order(row_ptr, col_idx, values, x):
    row_ptr = np.asarray(row_ptr, np.int64)
    n = len(row_ptr) - 1
    lens = np.diff(row_ptr)
    y = np.zeros(n)
    for k in range(int(lens.max()) if n else 0):
        rows = np.nonzero(lens > k)[0]
        o = row_ptr[rows] + k
        prod = values[o] * x[col_idx[o]]
        y[rows] = y[rows] + prod
    return y

// bitmap_blocks.hpp -- the reference's 4x4 bitmap-block matrix format:
// export (pack), import (unpack to CSR, on the host or the device) and a
// direct SpMV over it, so matrices in the layout of the reference's blocked
// FPGA SpMV run on this library and the library can hand its matrices to
// tools that consume that layout.
//
// Format (include/cg_common.hpp:26-50, packer
// archived/host/xrt_host.cpp:229-349):
//   * rows and columns in blocks of 4; block row br holds rows 4br..4br+3,
//     block column bc columns 4bc..4bc+3 (the last ones may pass n);
//   * a block is 16 FP64 values + 16 index bytes (144 bytes); bit
//     pos = 4 * local_row + local_col of the 16-bit mask
//     indices[2] | indices[3] << 8 marks a stored entry; the stored values
//     are compact at values[0..popcount), in ascending bit order; the other
//     values and index bytes are zero;
//   * the packer sums duplicate CSR entries of one position in CSR order
//     (starting from +0.0), keeps a position iff its sum != 0.0 and a block
//     iff its mask is nonzero, with block columns ascending in a block row;
//   * windowed layout (the reference's pcg_control_kernel input): columns in
//     windows of W (the reference's kSpmvWindowSize = 512), window w
//     covering block columns [floor(wW / 4), ceil(min(wW + W, n) / 4)); for
//     every window and block row, in that order, the block row's blocks of
//     the window's block columns; row_ptr holds num_windows * (nbr + 1)
//     entries, window w's at [w (nbr + 1), (w + 1)(nbr + 1));
//   * plain layout (W = 0 here; the standalone spmv_blocked_kernel input,
//     archived/kernels/spmv_blocked_kernel.cpp:131-161): one window over
//     every column, row_ptr of nbr + 1 entries.
// y = A x over either layout (archived/kernels/spmv_blocked_kernel.cpp:19-89
// and accumulate_window_tile, archived/kernels/pcg_control_kernel.cpp:
// 125-245): y[4br + r] starts at +0.0 and adds values * x for every stored
// entry of its row, windows in order, blocks in stored order, bits
// ascending, each product rounded and then added; an entry whose column is
// >= n (or, windowed, outside the window) multiplies 0.0.
#pragma once

#include <cstdint>
#include <vector>

#include "generators.hpp"

namespace xp {

constexpr int kBitmapSide = 4;
constexpr int kBitmapEntries = 16;

// Same layout as the reference's SpmvBlock (cg_common.hpp:47-50).
struct BitmapBlock {
    double values[kBitmapEntries];
    unsigned char indices[kBitmapEntries];
};
static_assert(sizeof(BitmapBlock) == 144, "SpmvBlock layout");

inline unsigned bitmap_mask(const BitmapBlock& b) {
    return static_cast<unsigned>(b.indices[2]) | static_cast<unsigned>(b.indices[3]) << 8;
}

inline int bitmap_block_rows(int n) { return (n + kBitmapSide - 1) / kBitmapSide; }
// Windows of a layout: window_size 0 is the plain layout (one window).
inline int bitmap_windows(int n, int window_size) {
    return window_size > 0 ? (n + window_size - 1) / window_size : 1;
}

struct BitmapMatrix {
    int n = 0;
    int window_size = 0;
    int num_windows = 1;
    int num_block_rows = 0;
    std::vector<int> row_ptr;          // num_windows * (num_block_rows + 1)
    std::vector<int> col_idx;          // block column of every block
    std::vector<BitmapBlock> blocks;
};

// The packer (window_size 0: plain layout).  Throws for a malformed CSR
// (the host builder's messages), window_size < 0 or more than INT_MAX
// blocks.
BitmapMatrix bitmap_pack(int n, const int* row_ptr, const int* col_idx, const double* values, int window_size);

// Checks a layout's pointers (throws "bitmap row_ptr ..." / "bitmap col_idx
// ..."): every window's pointers nondecreasing and within [0, nblocks],
// each window starting where the previous one ends (as the packer writes
// them), block columns not negative.
void bitmap_check(int n, int window_size, const int* row_ptr, const int* col_idx, long long nblocks);

// CSR of the matrix the layout describes: per row, the entries that y = A x
// reads, in its order (windows, blocks, bits), columns in [0, n) (entries
// outside the matrix or their window are dropped: they multiply 0.0);
// stored zeros are kept.  b = x0 = 0.
HostCsr bitmap_unpack(int n, int window_size, const int* row_ptr, const int* col_idx, const BitmapBlock* blocks,
                      long long nblocks);

}  // namespace xp

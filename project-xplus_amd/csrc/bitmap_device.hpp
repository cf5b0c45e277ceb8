// bitmap_device.hpp -- the 4x4 bitmap-block format on the device: y = A x
// straight from the block stream, and its import into a device CSR (from
// which the device set-up builds the solver's SELL matrix).  Format and
// arithmetic as described in bitmap_blocks.hpp.
#pragma once

#include "bitmap_blocks.hpp"
#include "hip_util.hpp"

namespace xp {

// Threads per workgroup of the bitmap kernels (one thread per matrix row).
constexpr int kBitmapThreads = 256;

// y[0..n) = A x (FP64; device pointers).  Pointers outside [0, nblocks] are
// clamped and negative block columns read x as 0.0, so a malformed layout
// cannot make the kernel read out of bounds (xp_bitmap_check-style
// validation is the importer's job).
void launch_bitmap_spmv(int n, int window_size, const int* row_ptr, const int* col_idx, const BitmapBlock* blocks,
                        long long nblocks, const double* x, double* y, hipStream_t s);

// Device CSR of the layout (bitmap_unpack's, entry for entry): validates
// the layout on the device (bitmap_check's messages), then counts, scans
// and fills.  Holds set_up_mutex while it uses the set-up's work buffers.
// nblocks <= kBitmapMaxBlocks keeps every count in int range.
constexpr long long kBitmapMaxBlocks = 134217727;  // INT_MAX / 16
void device_bitmap_csr(int n, int window_size, const int* row_ptr, const int* col_idx, const BitmapBlock* blocks,
                       long long nblocks, DevBuf<int>& rp, DevBuf<int>& ci, DevBuf<double>& va, hipStream_t s);

}  // namespace xp

// bitmap_device.hip -- kernels over the 4x4 bitmap-block format: the
// direct SpMV and the import into a device CSR.  See bitmap_device.hpp.
//
// One thread per matrix row: the four rows of a block row are four
// neighbouring lanes, so they read each block's column and mask from the
// same addresses (one request) and their value runs from the same 144-byte
// block, and a wavefront covers 16 block rows.  Every row sums its own
// entries sequentially in the format's order with separate FP64 rounding
// of product and sum (no fused multiply-add), which makes y bit for bit
// equal to the reference's sequential loops over the same blocks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <mutex>
#include <stdexcept>

#include "bitmap_device.hpp"
#include "device_build.hpp"

namespace xp {

namespace {

constexpr int kErrRange = 1, kErrOrder = 2, kErrCol = 4, kErrWindows = 8;

int groups_of(long long n) { return static_cast<int>((n + kBitmapThreads - 1) / kBitmapThreads); }

// The layout's geometry, shared by the kernels.
struct BitmapView {
    int n, nbr, nw, window;   // window 0: plain layout
    const int* row_ptr;
    const int* col_idx;
    const BitmapBlock* blocks;
    long long nblocks;
};

// The 16-bit mask of block k (bytes 130..131 of the block: 2-aligned).
__device__ __forceinline__ unsigned block_mask(const BitmapBlock* b) {
    return static_cast<unsigned>(*reinterpret_cast<const unsigned short*>(b->indices + 2));
}

// Calls f(column, value) for every entry of `row` that y = A x reads, in the
// format's order; entries outside the matrix or their window are passed
// with column -1 (they multiply 0.0).  Returns a bitmask of layout errors
// seen on the way (pointers out of [0, nblocks] or decreasing: the range
// read is clamped).
template <class F>
__device__ __forceinline__ int for_row_entries(const BitmapView& v, int row, F&& f) {
    const int br = row >> 2, r = row & 3;
    const unsigned below = (1u << (4 * r)) - 1u;
    int err = 0;
    for (int w = 0; w < v.nw; ++w) {
        const long long base = static_cast<long long>(w) * (v.nbr + 1) + br;
        long long lo = v.row_ptr[base], hi = v.row_ptr[base + 1];
        if (lo < 0 || hi > v.nblocks || lo > v.nblocks || hi < 0) err |= kErrRange;
        if (lo > hi) err |= kErrOrder;
        lo = max(lo, 0LL);
        hi = min(hi, v.nblocks);
        const long long wb = v.window > 0 ? static_cast<long long>(w) * v.window : 0;
        const long long we = v.window > 0 ? min(wb + v.window, static_cast<long long>(v.n)) : v.n;
        for (long long k = lo; k < hi; ++k) {
            const BitmapBlock* b = v.blocks + k;
            const unsigned mask = block_mask(b);
            const unsigned nib = (mask >> (4 * r)) & 0xFu;
            if (nib == 0) continue;
            int vi = __popc(mask & below);
            const long long c0 = static_cast<long long>(v.col_idx[k]) * 4;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                if ((nib >> c) & 1u) {
                    const long long gc = c0 + c;
                    f(gc >= wb && gc < we ? static_cast<int>(gc) : -1, b->values[vi]);
                    ++vi;
                }
            }
        }
    }
    return err;
}

// a * x rounded on its own (kept apart from the sum that follows: no fused
// multiply-add).
__device__ __forceinline__ double product(double a, double x) {
#pragma clang fp contract(off)
    return a * x;
}

__device__ __forceinline__ double sum(double acc, double p) {
#pragma clang fp contract(off)
    return acc + p;
}

// y = A x.  The same entries in the same order as for_row_entries, but the
// loads of kSpmvBatch blocks are issued together: first their masks and
// columns, then all their values and x entries, then the products are
// added in order.  A row's blocks then cost about two memory round trips
// per batch instead of two per block.  Measured on the thermal-FEM 1100^2
// matrix (2.1 M blocks, 4.0 entries each; the format streams 332 MB):
//   one block at a time                                   103 us
//   batches of 8 (182 VGPRs)                               87 us
//   batches of 4                                         80.6 us, 541 MB read
//   the 4 lanes of a block row copy each batch into LDS
//   with 16-byte loads, then read masks and values there  93.8 us, 392 MB read
//   a workgroup's blocks staged in LDS behind barriers     122 us
// The re-reads of this form (541 MB against 332 MB: the 16 block-row
// streams of a wavefront outlive the L2 between a block's mask load and its
// value loads) cost less than the extra dependent LDS step that removes them.
constexpr int kSpmvBatch = 4;

__global__ __launch_bounds__(kBitmapThreads) void k_bitmap_spmv(BitmapView v, const double* __restrict__ x,
                                                               double* __restrict__ y) {
    const long long row = static_cast<long long>(blockIdx.x) * kBitmapThreads + threadIdx.x;
    if (row >= v.n) return;
    const int br = static_cast<int>(row >> 2), r = static_cast<int>(row & 3);
    const unsigned below = (1u << (4 * r)) - 1u;
    double acc = 0.0;
    for (int w = 0; w < v.nw; ++w) {
        const long long base = static_cast<long long>(w) * (v.nbr + 1) + br;
        const long long lo = max(static_cast<long long>(v.row_ptr[base]), 0LL);
        const long long hi = min(static_cast<long long>(v.row_ptr[base + 1]), v.nblocks);
        const long long wb = v.window > 0 ? static_cast<long long>(w) * v.window : 0;
        const long long we = v.window > 0 ? min(wb + v.window, static_cast<long long>(v.n)) : v.n;
        for (long long k0 = lo; k0 < hi; k0 += kSpmvBatch) {
            unsigned nib[kSpmvBatch], first[kSpmvBatch];
            long long c0[kSpmvBatch];
#pragma unroll
            for (int j = 0; j < kSpmvBatch; ++j) {
                const bool ok = k0 + j < hi;
                const unsigned m = ok ? block_mask(v.blocks + k0 + j) : 0u;
                nib[j] = (m >> (4 * r)) & 0xFu;
                first[j] = __popc(m & below);
                c0[j] = ok ? static_cast<long long>(v.col_idx[k0 + j]) * 4 : 0;
            }
            double p[kSpmvBatch][4];
#pragma unroll
            for (int j = 0; j < kSpmvBatch; ++j) {
                const double* vals = v.blocks[k0 + j].values;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const bool has = (nib[j] >> c) & 1u;
                    const long long gc = c0[j] + c;
                    const double a = has ? vals[first[j] + __popc(nib[j] & ((1u << c) - 1u))] : 0.0;
                    const double xv = has && gc >= wb && gc < we ? x[gc] : 0.0;
                    p[j][c] = product(a, xv);
                }
            }
#pragma unroll
            for (int j = 0; j < kSpmvBatch; ++j)
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    if ((nib[j] >> c) & 1u) acc = sum(acc, p[j][c]);
        }
    }
    y[row] = acc;
}

__global__ __launch_bounds__(kBitmapThreads) void k_bitmap_count(BitmapView v, int* __restrict__ cnt,
                                                                int* __restrict__ err) {
    const long long row = static_cast<long long>(blockIdx.x) * kBitmapThreads + threadIdx.x;
    if (row >= v.n) return;
    int c = 0;
    const int e = for_row_entries(v, static_cast<int>(row), [&](int col, double) { c += col >= 0 ? 1 : 0; });
    cnt[row] = c;
    // windows follow each other: window w ends where window w + 1 starts
    // (which also bounds the entries of all windows together by 16 nblocks)
    int ew = 0;
    if (row == static_cast<long long>(v.nbr - 1) * 4)
        for (int w = 0; w + 1 < v.nw; ++w)
            if (v.row_ptr[static_cast<long long>(w) * (v.nbr + 1) + v.nbr] !=
                v.row_ptr[static_cast<long long>(w + 1) * (v.nbr + 1)])
                ew = kErrWindows;
    if (e | ew) atomicOr(err, e | ew);
}

__global__ __launch_bounds__(kBitmapThreads) void k_bitmap_cols(long long nblocks, const int* __restrict__ col_idx,
                                                               int* __restrict__ err) {
    const long long k = static_cast<long long>(blockIdx.x) * kBitmapThreads + threadIdx.x;
    if (k < nblocks && col_idx[k] < 0) atomicOr(err, kErrCol);
}

__global__ __launch_bounds__(kBitmapThreads) void k_bitmap_fill(BitmapView v, const int* __restrict__ rp,
                                                               int* __restrict__ ci, double* __restrict__ va) {
    const long long row = static_cast<long long>(blockIdx.x) * kBitmapThreads + threadIdx.x;
    if (row >= v.n) return;
    int o = rp[row];
    for_row_entries(v, static_cast<int>(row), [&](int col, double a) {
        if (col >= 0) {
            ci[o] = col;
            va[o] = a;
            ++o;
        }
    });
}

__global__ void k_bitmap_last(int n, const int* __restrict__ total, int* __restrict__ rp) {
    if (blockIdx.x == 0 && threadIdx.x == 0) rp[n] = *total;
}

BitmapView make_view(int n, int window_size, const int* row_ptr, const int* col_idx, const BitmapBlock* blocks,
                     long long nblocks) {
    if (n < 0) throw std::runtime_error("n must be non-negative");
    if (window_size < 0) throw std::runtime_error("window_size must be non-negative (0: plain layout)");
    if (nblocks < 0) throw std::runtime_error("nblocks must be non-negative");
    if (n > 0 && row_ptr == nullptr) throw std::runtime_error("null bitmap row_ptr");
    if (nblocks > 0 && (col_idx == nullptr || blocks == nullptr)) throw std::runtime_error("null bitmap array");
    BitmapView v{};
    v.n = n;
    v.nbr = bitmap_block_rows(n);
    v.nw = n > 0 ? bitmap_windows(n, window_size) : 0;
    v.window = window_size;
    v.row_ptr = row_ptr;
    v.col_idx = col_idx;
    v.blocks = blocks;
    v.nblocks = nblocks;
    return v;
}

}  // namespace

void launch_bitmap_spmv(int n, int window_size, const int* row_ptr, const int* col_idx, const BitmapBlock* blocks,
                        long long nblocks, const double* x, double* y, hipStream_t s) {
    const BitmapView v = make_view(n, window_size, row_ptr, col_idx, blocks, nblocks);
    if (n == 0) return;
    if (x == nullptr || y == nullptr) throw std::runtime_error("null vector");
    hipLaunchKernelGGL(k_bitmap_spmv, dim3(groups_of(n)), dim3(kBitmapThreads), 0, s, v, x, y);
    HIPX(hipGetLastError());
}

void device_bitmap_csr(int n, int window_size, const int* row_ptr, const int* col_idx, const BitmapBlock* blocks,
                       long long nblocks, DevBuf<int>& rp, DevBuf<int>& ci, DevBuf<double>& va, hipStream_t s) {
    const BitmapView v = make_view(n, window_size, row_ptr, col_idx, blocks, nblocks);
    if (nblocks > kBitmapMaxBlocks) throw std::runtime_error("more than 134,217,727 bitmap blocks");
    rp.alloc(static_cast<size_t>(n) + 1);
    DevBuf<int> cnt(std::max(n, 1)), flags(2);   // flags[0] errors, flags[1] total entries
    HIPX(hipMemsetAsync(flags.p, 0, 2 * sizeof(int), s));
    if (nblocks > 0)
        hipLaunchKernelGGL(k_bitmap_cols, dim3(groups_of(nblocks)), dim3(kBitmapThreads), 0, s, nblocks, col_idx,
                           flags.p);
    if (n > 0) hipLaunchKernelGGL(k_bitmap_count, dim3(groups_of(n)), dim3(kBitmapThreads), 0, s, v, cnt.p, flags.p);
    HIPX(hipGetLastError());
    {
        std::lock_guard<std::mutex> lock(set_up_mutex());
        device_scan_exclusive(cnt.p, rp.p, n, flags.p + 1, s);
    }
    hipLaunchKernelGGL(k_bitmap_last, dim3(1), dim3(64), 0, s, n, flags.p + 1, rp.p);
    int h[2] = {0, 0};
    HIPX(hipMemcpyAsync(h, flags.p, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPX(hipStreamSynchronize(s));
    if (h[0] & kErrRange) throw std::runtime_error("bitmap row_ptr out of range");
    if (h[0] & kErrOrder) throw std::runtime_error("bitmap row_ptr must be nondecreasing");
    if (h[0] & kErrWindows) throw std::runtime_error("bitmap row_ptr windows must be contiguous");
    if (h[0] & kErrCol) throw std::runtime_error("bitmap col_idx must be non-negative");
    const size_t nnz = static_cast<size_t>(h[1]);
    ci.alloc(std::max<size_t>(nnz, 1));
    va.alloc(std::max<size_t>(nnz, 1));
    if (n > 0) hipLaunchKernelGGL(k_bitmap_fill, dim3(groups_of(n)), dim3(kBitmapThreads), 0, s, v, rp.p, ci.p, va.p);
    HIPX(hipGetLastError());
}

}  // namespace xp

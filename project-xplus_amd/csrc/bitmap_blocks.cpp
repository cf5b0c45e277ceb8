// bitmap_blocks.cpp -- 4x4 bitmap-block format: pack / check / unpack.
#include "bitmap_blocks.hpp"

#include <algorithm>
#include <climits>
#include <stdexcept>
#include <string>

namespace xp {

namespace {

// One stored position of a block row before packing.
struct Entry {
    int bc;
    int pos;
    double v;
};

void check_csr(int n, const int* row_ptr, const int* col_idx) {
    if (n < 0) throw std::runtime_error("n must be non-negative");
    if (row_ptr[0] != 0) throw std::runtime_error("row_ptr[0] must be 0");
    for (int i = 0; i < n; ++i) {
        if (row_ptr[i] > row_ptr[i + 1]) throw std::runtime_error("row_ptr must be nondecreasing");
        for (int o = row_ptr[i]; o < row_ptr[i + 1]; ++o)
            if (col_idx[o] < 0 || col_idx[o] >= n) throw std::runtime_error("col_idx out of range");
    }
}

}  // namespace

BitmapMatrix bitmap_pack(int n, const int* row_ptr, const int* col_idx, const double* values, int window_size) {
    if (window_size < 0) throw std::runtime_error("window_size must be non-negative (0: plain layout)");
    check_csr(n, row_ptr, col_idx);
    BitmapMatrix out;
    out.n = n;
    out.window_size = window_size;
    out.num_block_rows = bitmap_block_rows(n);
    out.num_windows = bitmap_windows(n, window_size);
    const int nbr = out.num_block_rows;

    // Blocks of every block row, block columns ascending (row_first[br] ..).
    std::vector<long long> row_first(static_cast<size_t>(nbr) + 1, 0);
    std::vector<int> bcol;
    std::vector<BitmapBlock> blk;
    std::vector<Entry> ents;
    for (int br = 0; br < nbr; ++br) {
        ents.clear();
        for (int r = 0; r < kBitmapSide; ++r) {
            const int row = br * kBitmapSide + r;
            if (row >= n) break;
            for (int o = row_ptr[row]; o < row_ptr[row + 1]; ++o)
                ents.push_back({col_idx[o] / kBitmapSide, r * kBitmapSide + col_idx[o] % kBitmapSide, values[o]});
        }
        // stable: entries of one position keep their CSR order for the sum
        std::stable_sort(ents.begin(), ents.end(), [](const Entry& a, const Entry& b) { return a.bc < b.bc; });
        for (size_t k = 0; k < ents.size();) {
            const int bc = ents[k].bc;
            double dense[kBitmapEntries] = {};
            for (; k < ents.size() && ents[k].bc == bc; ++k) dense[ents[k].pos] += ents[k].v;
            BitmapBlock b{};
            unsigned mask = 0;
            int nv = 0;
            for (int pos = 0; pos < kBitmapEntries; ++pos) {
                if (dense[pos] != 0.0) {
                    mask |= 1u << pos;
                    b.values[nv++] = dense[pos];
                }
            }
            if (mask == 0) continue;
            b.indices[2] = static_cast<unsigned char>(mask & 0xFFu);
            b.indices[3] = static_cast<unsigned char>(mask >> 8);
            bcol.push_back(bc);
            blk.push_back(b);
        }
        row_first[static_cast<size_t>(br) + 1] = static_cast<long long>(blk.size());
    }

    if (window_size == 0) {
        if (blk.size() > static_cast<size_t>(INT_MAX)) throw std::runtime_error("more than INT_MAX blocks");
        out.row_ptr.resize(static_cast<size_t>(nbr) + 1);
        for (int br = 0; br <= nbr; ++br) out.row_ptr[static_cast<size_t>(br)] = static_cast<int>(row_first[br]);
        out.col_idx = std::move(bcol);
        out.blocks = std::move(blk);
        return out;
    }
    // Window-major arrangement: for every window, every block row's blocks
    // whose column lies in the window's block-column range.  A window size
    // that is not a multiple of 4 makes neighbouring ranges share one block
    // column, whose block is then listed in both windows.
    const int nw = out.num_windows;
    out.row_ptr.reserve(static_cast<size_t>(nw) * (static_cast<size_t>(nbr) + 1));
    for (int w = 0; w < nw; ++w) {
        const long long wb = static_cast<long long>(w) * window_size;
        const long long we = std::min<long long>(wb + window_size, n);
        const int bc_lo = static_cast<int>(wb / kBitmapSide);
        const int bc_hi = static_cast<int>((we + kBitmapSide - 1) / kBitmapSide);
        for (int br = 0; br < nbr; ++br) {
            if (out.blocks.size() > static_cast<size_t>(INT_MAX)) throw std::runtime_error("more than INT_MAX blocks");
            out.row_ptr.push_back(static_cast<int>(out.blocks.size()));
            const auto first = bcol.begin() + row_first[br], last = bcol.begin() + row_first[br + 1];
            auto it = std::lower_bound(first, last, bc_lo);
            for (; it != last && *it < bc_hi; ++it) {
                const size_t k = static_cast<size_t>(it - bcol.begin());
                out.col_idx.push_back(bcol[k]);
                out.blocks.push_back(blk[k]);
            }
        }
        if (out.blocks.size() > static_cast<size_t>(INT_MAX)) throw std::runtime_error("more than INT_MAX blocks");
        out.row_ptr.push_back(static_cast<int>(out.blocks.size()));
    }
    return out;
}

void bitmap_check(int n, int window_size, const int* row_ptr, const int* col_idx, long long nblocks) {
    if (n < 0) throw std::runtime_error("n must be non-negative");
    if (window_size < 0) throw std::runtime_error("window_size must be non-negative (0: plain layout)");
    if (nblocks < 0) throw std::runtime_error("nblocks must be non-negative");
    const int nbr = bitmap_block_rows(n), nw = n > 0 ? bitmap_windows(n, window_size) : 0;
    for (int w = 0; w < nw; ++w) {
        const int* p = row_ptr + static_cast<size_t>(w) * (static_cast<size_t>(nbr) + 1);
        for (int br = 0; br <= nbr; ++br) {
            if (p[br] < 0 || p[br] > nblocks) throw std::runtime_error("bitmap row_ptr out of range");
            if (br < nbr && p[br] > p[br + 1]) throw std::runtime_error("bitmap row_ptr must be nondecreasing");
        }
    }
    // windows follow each other, so all windows together list at most
    // nblocks blocks
    for (int w = 0; w + 1 < nw; ++w) {
        const size_t end = static_cast<size_t>(w) * (static_cast<size_t>(nbr) + 1) + static_cast<size_t>(nbr);
        if (row_ptr[end] != row_ptr[end + 1]) throw std::runtime_error("bitmap row_ptr windows must be contiguous");
    }
    for (long long k = 0; k < nblocks; ++k)
        if (col_idx[k] < 0) throw std::runtime_error("bitmap col_idx must be non-negative");
}

HostCsr bitmap_unpack(int n, int window_size, const int* row_ptr, const int* col_idx, const BitmapBlock* blocks,
                      long long nblocks) {
    bitmap_check(n, window_size, row_ptr, col_idx, nblocks);
    HostCsr a;
    a.n = n;
    a.row_ptr.assign(static_cast<size_t>(n) + 1, 0);
    a.b.assign(static_cast<size_t>(n), 0.0);
    a.x0.assign(static_cast<size_t>(n), 0.0);
    const int nbr = bitmap_block_rows(n), nw = bitmap_windows(n, window_size);
    for (int row = 0; row < n; ++row) {
        const int br = row / kBitmapSide, r = row % kBitmapSide;
        for (int w = 0; w < nw; ++w) {
            const long long wb = window_size > 0 ? static_cast<long long>(w) * window_size : 0;
            const long long we = window_size > 0 ? std::min<long long>(wb + window_size, n) : n;
            const int* p = row_ptr + static_cast<size_t>(w) * (static_cast<size_t>(nbr) + 1);
            for (int k = p[br]; k < p[br + 1]; ++k) {
                const unsigned mask = bitmap_mask(blocks[k]);
                const unsigned nib = (mask >> (kBitmapSide * r)) & 0xFu;
                int vi = __builtin_popcount(mask & ((1u << (kBitmapSide * r)) - 1u));
                for (int c = 0; c < kBitmapSide; ++c) {
                    if (!((nib >> c) & 1u)) continue;
                    const long long gc = static_cast<long long>(col_idx[k]) * kBitmapSide + c;
                    if (gc >= wb && gc < we) {
                        a.col_idx.push_back(static_cast<int>(gc));
                        a.values.push_back(blocks[k].values[vi]);
                    }
                    ++vi;
                }
            }
        }
        if (a.col_idx.size() > static_cast<size_t>(INT_MAX)) throw std::runtime_error("more than INT_MAX entries");
        a.row_ptr[static_cast<size_t>(row) + 1] = static_cast<int>(a.col_idx.size());
    }
    return a;
}

}  // namespace xp

// sell_build.hpp -- host-side CSR -> SELL-C-sigma conversion (no GPU code).
//
// The SELL matrix replaces the reference's Cuper slice/window packing
// (host/cuper_pcg_solver.hpp:90-134 build_cuper_matrix_plan,
// host/cuper_control_matrix.hpp:130-207 build_cuper_control_matrix): instead of
// 8192-column x batches streamed through FPGA channels, rows are grouped into
// 64-row slices (one per wavefront) with the diagonal stored first, and the
// column indices compressed to 16-bit deltas where the slice span allows.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "blocking.hpp"
#include "xp_types.hpp"

namespace xp {

struct CsrRef {
    int n = 0;
    const int* row_ptr = nullptr;
    const int* col_idx = nullptr;
    const double* values = nullptr;
};

struct SellHost {
    int n = 0;            // real rows over all systems
    int n_pad = 0;        // padded rows (multiple of kChunkRows per system)
    int nslices = 0;
    int nsys = 0;
    int sigma = 0;
    int idx16 = 0;
    int block_rows = 1;
    long long nnz = 0;            // real stored entries
    long long stored = 0;         // slice_ptr[nslices] * 64, including padding
    std::vector<int> perm;        // solve row -> natural global row, -1 for padding
    std::vector<int> iperm;       // natural global row -> solve row
    std::vector<int> sys_row_begin;  // [nsys + 1] solve-order row offsets (padded)
    std::vector<int> sys_nat_begin;  // [nsys + 1] natural row offsets
    std::vector<int> slice_ptr;
    std::vector<int> slice_base;
    std::vector<float> val;
    std::vector<float> val_lo;
    std::vector<int16_t> val_lo16;  // with keep_lo: the same remainder in 16 bits (see lo16_of)
    std::vector<double> val64;
    std::vector<uint16_t> col16;
    std::vector<int> col32;
    std::vector<uint8_t> has_diag;  // per solve row: 1 if the CSR row had a diagonal entry
    // block_rows = 8: block (unit) id of every 8-row solve block (units of
    // all systems numbered consecutively), -1 for an all-padding block
    std::vector<int> block_unit;
    std::vector<int> unit_first;    // [nsys + 1] first unit id of each system
};

// Builds the SELL form of a block-diagonal batch of CSR systems.  sigma is
// the sort window in rows (a multiple of 64, or 0/64 for no sorting).
// keep_lo / keep64 request the hi+lo and exact-FP64 value streams.
// Throws std::runtime_error on malformed input.
// block_rows = 8 keeps each block of rows together as 8 consecutive solve
// rows (sorted as units; smaller blocks padded), for block preconditioners:
// the blocks of `blocks` (one Blocking per system), or 8 consecutive natural
// rows when it is null.
// The remainder v - (double)hi of an FP64 value v with FP32 part hi, as an
// integer count of 2^(E-38) where hi = m 2^E, 1 <= |m| < 2: |remainder| <=
// ulp(hi) / 2 = 2^(E-24), so the count fits in 15 bits, and hi + count *
// 2^(E-38) reproduces v to ~2^-39 relative.  0 for hi = 0.
int16_t lo16_of(double v, float hi);

SellHost build_sell(const std::vector<CsrRef>& systems, int sigma, bool keep_lo, bool keep64,
                    int block_rows = 1, const std::vector<Blocking>* blocks = nullptr);

}  // namespace xp

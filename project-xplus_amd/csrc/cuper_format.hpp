// cuper_format.hpp -- the reference's Cuper packed matrix format: export
// (pack) and import (unpack), so matrices in the reference's FPGA layout run
// on this library's SpMV and the library can hand matrices to tools that
// consume that layout.
//
// Format (host/cuper_control_matrix.hpp:14-207, the word layout also of
// DLC/Cuper/include/Cuper_common.h:535-774):
//   * columns in batches of 8192; rows dealt to 128 physical PEs
//     (16 HBM channels x 8 lanes) by cuper_pe_for_row();
//   * per batch and PE the entries are sorted by column (std::sort on the
//     column, entries pushed in CSR order), then scheduled so that entries
//     of one row group (row / 256) sit at least 10 slots apart, and every
//     PE list is padded to the batch's longest; sp_element_list_ptr[b + 1]
//     is the padded length after batch b;
//   * channel c holds, for index i, the 8 words of PEs 8c .. 8c+7 at
//     [8 i, 8 i + 8); channel length rounded up to 512 words;
//   * word = local column (14 bits) << 50 | row code (18 bits) << 32 |
//     FP32 value bits; row code = (row / 256) * 2 + (row & 1); an empty slot
//     is 0x3FFFF << 32 (bit 17 of the row code set).
// pack reproduces the reference's packer word for word (the column sort is
// the same std::sort over the same input order, so ties land where the
// reference's land with the same C++ standard library).  unpack inverts the
// row mapping as the kernel's restore_global_row does
// (kernels/cuper_pcg_control_kernel.cpp:140-147).
#pragma once

#include <cstdint>
#include <vector>

#include "generators.hpp"

namespace xp {

constexpr int kCuperSliceWidth = 8192;
constexpr int kCuperChannels = 16;
constexpr int kCuperLanes = 8;
constexpr int kCuperPes = kCuperChannels * kCuperLanes;
constexpr int kCuperWindow = 10;
constexpr long long kCuperMaxRows = 1LL << 24;   // row code must leave bit 17 clear

struct CuperPacked {
    int batch_count = 0;
    int matrix_len = 0;        // padded entries per PE list (= ptr.back())
    int channel_size = 0;      // words per channel (multiple of 512)
    std::vector<int> ptr;      // [batch_count + 1]
    std::vector<uint64_t> data;  // [kCuperChannels * channel_size], channel-major
};

int cuper_pe_for_row(int row);
// n rows; values rounded to FP32.  Throws for n > kCuperMaxRows.
CuperPacked cuper_pack(int n, const int* row_ptr, const int* col_idx, const double* values);
// CSR of an n x n matrix (rows ascending, columns ascending within a row,
// entries of one (row, column) kept in slot order), values widened from
// FP32; b = x0 = 0.  Throws on entries outside the matrix or a malformed
// batch table.
HostCsr cuper_unpack(const CuperPacked& p, int n);

}  // namespace xp

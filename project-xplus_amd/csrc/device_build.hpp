// device_build.hpp -- the matrix set-up on the device: CSR (resident in
// HBM) -> strength aggregates of up to 8 rows -> SELL-C-sigma in solve order.
//
// The host builder (sell_build.cpp) and the host aggregation rules
// (blocking.cpp) read the CSR on the host, which puts a PCIe copy and a
// serial pass over every nonzero in front of each solve; here the same
// SELL arrays are produced by kernels from a device CSR, bit for bit equal
// to the host builder's for the same blocks, and the blocks come from a
// parallel form of the strength-ordered rule (blocking rule 3):
//   rounds: seeds are the unassigned rows that precede every unassigned
//           row within distance 2 (larger total strength t_i, then smaller
//           row); then kGrowSteps growth steps in which every unassigned
//           row bids for the neighbouring block with room that holds the
//           largest fraction of its strength (ties: smaller block id) and
//           every block takes its best bid (FP32-rounded fraction, ties:
//           smaller row); rounds repeat until every row has a block;
//   regroup: rows of blocks smaller than 8 are pooled and the rounds rerun
//           on the pool;
//   pack:   the remaining pieces, in order of their smallest row, in runs
//           of kPackSegment pieces (one thread each): each piece into the
//           most recent of the run's last kPackWindow packed blocks with
//           room, else a new one;
//   rows ascending inside a block, blocks ordered by their smallest row.
// On the thermal-FEM family its blocks are as good as the sequential rule
// 2's (exact block inverses: 1,328 against 1,339 PCG iterations at 202^2,
// 2,656 against 2,642 at 402^2).
#pragma once

#include <cstdint>
#include <mutex>

#include "hip_util.hpp"
#include "xp_types.hpp"

namespace xp {

constexpr int kGrowSteps = 8;
constexpr int kPackWindow = 16;
constexpr int kPackSegment = 32;

// A CSR system in device memory (one system).
struct DevCsr {
    int n = 0;
    long long nnz = 0;
    const int* row_ptr = nullptr;
    const int* col_idx = nullptr;
    const double* values = nullptr;
};

// Per-row facts: position of the first diagonal entry (-1 if missing),
// off-diagonal count (every entry but that one), D^-1/2 from the first
// diagonal entry (1 where it is missing or not positive) and the total
// off-diagonal strength t_i = sum |a_ij| d_i d_j (CSR order).
// (The arrays are the set-up's reusable work buffers: valid until the next
// build; builds are serialised by set_up_mutex.)
struct DevRows {
    int* offcnt = nullptr;
    int* diag_pos = nullptr;
    double* dis = nullptr;
    double* total = nullptr;
};

std::mutex& set_up_mutex();

// Validates the CSR on the device (throws the host builder's messages:
// "row_ptr[0] must be 0", "row_ptr must be nondecreasing", "col_idx out of
// range") and fills `rows`; sets a.nnz from row_ptr[n].
// trusted_nnz >= 0: a CSR this library produced with that many entries
// (the multilevel hierarchy's Galerkin products): the checks and their two
// host round trips are skipped.
void device_rows(DevCsr& a, DevRows& rows, hipStream_t s, int trusted_nnz = -1);

struct AggStats {
    int rounds = 0;          // seed rounds over all rows
    int regroup_rounds = 0;  // rounds over the pooled rows
    int pieces = 0;          // blocks smaller than 8 after the regroup
    int units = 0;           // final blocks
};

// Rule-3 aggregation: slots[8u + j] = row of slot j of block u (-1 for
// padding, real rows first, ascending).  Returns the block count.
// seed_dist 1: seeds need only precede their distance-1 neighbours (rule 4,
// for the dense coarse levels of the multilevel hierarchy).
int device_aggregate(const DevCsr& a, const DevRows& rows, DevBuf<int>& slots, AggStats* st, hipStream_t s,
                     int seed_dist = 2);

struct DevSell {
    int n = 0, n_pad = 0, nslices = 0, sigma = 0, idx16 = 0, max_width = 0, units = 0;
    long long nnz = 0, stored = 0;
    DevBuf<int> perm, iperm, slice_ptr, slice_base, block_unit;
    DevBuf<float> val, val_lo;
    DevBuf<int16_t> val_lo16;
    DevBuf<double> val64;
    DevBuf<uint16_t> col16;
    DevBuf<int> col32;
    DevBuf<double> diag;   // natural order, the first diagonal entry (0 if missing)
    DevBuf<float> minv;    // solve order, 1 / diag (padding rows 1)
    int bad_row = -1;      // first natural row whose diagonal fails the Jacobi checks
    int bad_kind = 0;      // 1 missing, 2 |d| <= DBL_MIN, 3 |d| <= 1e-30
};

// SELL-C-sigma of one system from its device CSR: block_rows = 1 (rows
// sorted by off-diagonal count inside windows of sigma rows) or 8 (the
// `units` blocks of d_slots kept as 8 consecutive solve rows, sorted as
// units by their longest row).  keep_lo / keep64: the hi + lo and FP64
// value streams.  Equal, array for array, to xp::build_sell of the same
// CSR and blocks.
void device_sell(const DevCsr& a, const DevRows& rows, int sigma, bool keep_lo, bool keep64, int block_rows,
                 const int* d_slots, int units, DevSell& out, hipStream_t s);

// Blocks of 8 consecutive rows (the last one padded): ceil(n / 8) units.
int device_natural_slots(int n, DevBuf<int>& slots, hipStream_t s);

// out[i] = sum in[0..i), *total (device pointer, may be null) = sum in[0..n),
// n >= 0.  Uses the set-up's work buffers: call with set_up_mutex held.
void device_scan_exclusive(const int* in, int* out, int n, int* total, hipStream_t s);

// out[k] = sum of x[i]^2 over the natural rows [sys_begin[k], sys_begin[k+1])
// of system k, in a fixed order; part: kSysSumParts * nsys doubles of scratch.
constexpr int kSysSumParts = 64;
void launch_sys_sumsq(int nsys, const int* sys_begin, const double* x, double* out, double* part, hipStream_t s);

}  // namespace xp

// xp_types.hpp -- data layout shared by the host builder and the gfx950 kernels.
//
// Matrix format: SELL-C-sigma with C = 64 (one wavefront lane per row, one
// wavefront per slice), rows sorted by length inside windows of sigma rows.
// The sort is applied to the whole system as a symmetric permutation, so the
// solver runs in "solve order": every vector is contiguous in that order and
// the SpMV writes y coalesced.  Vectors are padded to n_pad (a multiple of
// the work-chunk height); padding rows are identity rows with b = x0 = 0, so
// they stay exactly zero and add nothing to any dot product.
//
// Inside slice s (rows 64s .. 64s+63), column j of the slice occupies the 64
// consecutive entries [(slice_ptr[s] + j) * 64, +64): lane l holds entry j of
// row 64s+l.  Column j = 0 is the diagonal of every row (its column index is
// the row itself and is never stored or loaded); columns j >= 1 hold the
// off-diagonal entries, padded with value 0 and delta 0.  Off-diagonal column
// indices are stored as 16-bit deltas from a per-slice base column when every
// slice's column span fits, else as absolute 32-bit indices.
#pragma once

#include <cstdint>

namespace xp {

constexpr int kC = 64;               // slice height == wavefront width on CDNA
constexpr int kWaves = 16;           // wavefronts per workgroup in the streaming kernels
constexpr int kBlock = kC * kWaves;  // 1024 threads
constexpr int kNumXcd = 8;           // MI355X: 8 XCDs, workgroups dispatched round-robin

// One logical workgroup's share of the rows: slices [slice_begin, slice_end)
// of system `sys` (rows 64*slice_begin .. 64*slice_end-1).  A workgroup
// never straddles two systems; sys < 0 marks an idle workgroup.
struct WorkItem {
    int sys;
    int slice_begin;
    int slice_end;
    int pad_;
};

// Reduction slots of the per-workgroup partial sums (partials[slot * grid + block]).
// Partial slots.  pAp alternates between kSlotPap and kSlotPap1 by
// iteration parity (Launch::pap_slot): the fused kernel reads the previous
// iteration's pAp partials in its prologue while its own workgroups write
// the new ones, so the two must not share a slot.
enum : int { kSlotPap = 0, kSlotRz = 1, kSlotRr = 2, kSlotAux = 3, kSlotPap1 = 4, kNumSlots = 5 };

// Per-system PCG state machine, resident on the device.  Written only by the
// last workgroup of a reduction (see kernels.hip), read by every kernel of
// the next step.  Layout is plain-old-data and mirrored by xp_status_t.
struct SysState {
    double rz;
    double rr;
    double pap;
    double alpha;
    double beta;
    double tau;       // absolute bound on r^T r
    double rr0;       // r^T r after init
    double best_rr;   // smallest rr seen (stagnation watch)
    double rz_prev;   // rz before the last scalar step (residual replacement rewinds to it)
    int iters;
    int max_iters;
    int status;       // 0 converged, 1 max_iter, 2 breakdown, 3 stagnation, -1 running
    int active;       // 1 while iterating
    int stall;        // iterations since best_rr last improved by the stagnation factor
    int stall_limit;  // 0 = stagnation watch off
    int trace_cap;
    int fresh;        // 1 right after init (fused iteration: no previous step to finish)
    int x_pend;       // accurate iteration: x still lacks alpha p of the system's last
                      // step (it stopped after an even update: k_pcg_flush_x_acc adds it)
};

// Per-system Jacobi iteration state (two slots alternating per update,
// like SysState).  The stop test of update k runs at the start of the next
// kernel, from the partial maxima update k wrote.
struct JacState {
    double diff;      // max_i |x_new_i - x_old_i| of the last tested update
    double tau;       // with use_tau: converged when diff <= tau
    int iters;        // updates done
    int max_iters;
    int status;       // kStatus* below
    int active;       // 1 while iterating
    int pending;      // the last update's diff has not been tested yet
    int use_tau;      // 0: fixed count of max_iters updates
};

enum : int {
    kStatusConverged = 0,
    kStatusMaxIter = 1,
    kStatusBreakdown = 2,
    kStatusStagnation = 3,
    kStatusRunning = -1,
};

// One trace row per completed iteration (the reference's
// GoldenIterationTrace minus the derived residual = sqrt(rr)).
struct TraceRow {
    double alpha;
    double beta;
    double rz;
    double rr;
};

// A block preconditioner's block q (solve rows 8q .. 8q+7), factored:
// M_q = diag(s) C C^T diag(s) with C lower triangular with unit-scale
// entries (the Cholesky factor of M_q with its diagonal scaled to 1), the
// 36 entries stored as FP16 in four 10-entry pairs of rows: pair p (0..3)
// holds row p (p + 1 entries) then row 7 - p (8 - p entries) and one pad,
// so the lane of row i finds its whole row in the 20 bytes of pair
// min(i, 7 - i) (block_half_index).  112 bytes per block, 14 per row.
struct alignas(16) BlockPre {
    uint16_t c[40];
    float s[8];
};
static_assert(sizeof(BlockPre) == 112, "BlockPre layout");

// Position in BlockPre::c of C[i][k], k <= i.
constexpr int block_pair(int i) { return i < 4 ? i : 7 - i; }
constexpr int block_row_offset(int i) { return i < 4 ? 0 : 8 - i; }  // within the pair
constexpr int block_half_index(int i, int k) { return 10 * block_pair(i) + block_row_offset(i) + k; }

// Device view of a SELL matrix (all pointers are device pointers).
struct SellView {
    int n_pad;
    int nslices;
    int nsys;
    int idx16;               // 1: col16 deltas from slice_base; 0: col32 absolute
    int variant;             // SpMV inner loop: 0 rolled loop, 1 fixed-width unrolled
    int max_width;           // widest slice, in columns (diagonal included)
    const int* slice_ptr;    // [nslices + 1], in units of 64-entry columns
    const int* slice_base;   // [nslices]
    const float* val;        // [slice_ptr[nslices] * 64], diagonal first
    const float* val_lo;     // optional: v - (double)val, rounded to float
    const int16_t* val_lo16; // optional: v - (double)val in units of 2^(E-38), E = exponent of val
                             // (val + lo16 carries ~39 bits: the accurate iteration's SpMV)
    const double* val64;     // optional: exact FP64 values
    const uint16_t* col16;
    const int* col32;
};

}  // namespace xp

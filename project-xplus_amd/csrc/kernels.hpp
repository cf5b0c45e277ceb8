// kernels.hpp -- launchers for the gfx950 kernels in kernels.hip.
//
// Every launcher is stream-ordered and allocates nothing; buffers belong to
// the caller (the runtime in runtime.hip).  All pointers are device pointers.
#pragma once

#include <hip/hip_runtime.h>

#include "xp_types.hpp"

namespace xp {

struct MgPlan;  // the multilevel preconditioner's V-cycle (below)

// Shared launch context of the row-streaming kernels: the work map (one
// WorkItem per logical workgroup), the per-system workgroup ranges and the
// per-workgroup reduction partials (kNumSlots x grid doubles).
struct Launch {
    int grid = 0;                       // workgroups, a multiple of kNumXcd
    const WorkItem* work = nullptr;     // [grid]
    const int2* sys_blocks = nullptr;   // [nsys] logical block range per system
    double* partials = nullptr;         // [kNumSlots * grid]
    int u_blocks = 0;                   // > 0: one system, uniform split of u_slices over u_blocks
    int u_slices = 0;
    hipStream_t stream = nullptr;
    // pAp partial slots of the PCG iteration: this iteration's (pap_slot)
    // and the previous one's (pap_prev); set_parity alternates them.
    int pap_slot = kSlotPap, pap_prev = kSlotPap1;
    void set_parity(int k) {
        pap_slot = (k & 1) ? kSlotPap1 : kSlotPap;
        pap_prev = (k & 1) ? kSlotPap : kSlotPap1;
    }
    // Per-kernel timing only (never under graph capture): when set, the
    // iteration and SpMV launchers dispatch their kernel with these events
    // recorded at its start and end (hipExtLaunchKernelGGL).
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
};

// ---- FP32 solve-order PCG step kernels (state machine in SysState) ----
// init: r = b - A x; z = minv r; p = z; then the state after init in st[sys]
// (which must hold the configuration: tau, max_iters, trace_cap, stall_limit).
// blockM: null for Jacobi (z = minv r), else factored 8x8 blocks (z = M_b r_b).
void launch_pcg_init(const Launch& L, const SellView& A, SysState* st, const float* x,
                     const float* b, const float* minv, const BlockPre* blockM, float* r, float* z, float* p);
// ap = A p and the pAp partials.
void launch_pcg_spmv(const Launch& L, const SellView& A, const SysState* st, const float* p, float* ap);
// Fused iteration, first kernel: scalar tail of the previous iteration,
// p = z + beta p, ap = A z + beta ap, pAp partials; reads st_cur, writes st_next.
void launch_pcg_spmv_fused(const Launch& L, const SellView& A, const SysState* st_cur, SysState* st_next,
                           TraceRow* trace, int trace_stride, const float* z, float* p, float* ap);
// alpha; x += alpha p; r -= alpha ap; z = minv r; rz, rr partials.
void launch_pcg_update_xrz(const Launch& L, const SysState* st, float* x, const float* p, float* r,
                           const float* ap, const float* minv, const BlockPre* blockM, float* z);
// scalar tail (beta, tests, trace) and p = z + beta p; reads st_cur, writes st_next.
void launch_pcg_update_p(const Launch& L, const SysState* st_cur, SysState* st_next, TraceRow* trace,
                         int trace_stride, const float* z, float* p);

// ---- accurate iteration: FP32 r, z; FP64 p, ap, x; FP64-accumulated SpMV
// with hi+lo (use_lo) or FP64 values ----
// init: r = fp32(b - A x) (FP64), z = M r, rz / rr partials, then the state
// after init; the first fused iteration sets p = z, ap = A z (state.fresh).
void launch_pcg_init_acc(const Launch& L, const SellView& A, bool use_lo, SysState* st, const double* x,
                         const double* b, const float* minv, const BlockPre* blockM, float* r, float* z,
                         const MgPlan* mg = nullptr);
// vmode: 0 FP64 values, 1 hi + lo (FP32 pair), 2 hi + lo16.  p alternates
// between two buffers by iteration parity (p_old the previous iteration's).
void launch_pcg_spmv_fused_acc(const Launch& L, const SellView& A, int vmode, const SysState* st_cur,
                               SysState* st_next, TraceRow* trace, int trace_stride, const float* z,
                               const double* p_old, double* p_new, double* ap);
// odd iterations also add both steps' alpha p to x (p_prev: the previous
// iteration's direction); even ones leave x alone.
void launch_pcg_update_xrz_acc(const Launch& L, const SysState* st, bool odd, double* x, const double* p,
                               const double* p_prev, float* r, const double* ap, const float* minv,
                               const BlockPre* blockM, float* z);
// x += alpha p for the systems with state.x_pend (stopped behind an even
// update), then clears the flag; st: the current state slot.
void launch_pcg_flush_x_acc(const Launch& L, int nsys, SysState* st, const double* p, double* x);

// ---- multilevel (aggregation) preconditioner ----
constexpr int kMaxLevels = 16;
// One SELL level of the hierarchy (level 0: the solver's matrix).  Row u of
// level l + 1 is aggregate u of level l; mp / mpos list, for every solve row
// of level l >= 1, its aggregate's rows as solve positions of level l - 1.
struct MgLevelView {
    Launch L;                    // the level's launch context (level 0: unused, the solver's is passed)
    SellView A;
    const BlockPre* M = nullptr; // block factors of the smoother
    const int* perm = nullptr;   // solve row -> natural row, -1 for padding rows
    float* r = nullptr;          // level >= 1: restricted residual
    float* z = nullptr;          // level >= 1: the level's correction
    float* w = nullptr;          // the residual r - A z the restriction sums (every level)
    float* t = nullptr;          // prolonged approximation (every level)
    const int* mp = nullptr;     // level >= 1: [n_pad + 1] list offsets (padding rows: empty)
    const int* mpos = nullptr;   // level >= 1: the lists
};
// Where a kernel that produces a coarse correction e prolongs it: the finer
// level's t = cz z + gamma e on the rows of the coarse row's aggregate.
struct MgUp {
    const int* mp = nullptr;
    const int* mpos = nullptr;
    float* t = nullptr;
    const float* z = nullptr;
    float cz = 1.0f;
};
struct MgPlan {
    int nlev = 0;                // SELL levels; the dense coarsest level follows them
    MgLevelView lv[kMaxLevels];
    int nc = 0;                  // rows of the coarsest level (at most 4096)
    const float* cinv = nullptr; // its dense inverse, nc x nc row-major FP32
    const int* cmp = nullptr;    // its rows' aggregates in the last SELL level (as MgLevelView::mp / mpos)
    const int* cmpos = nullptr;
    float* ec = nullptr;         // its correction
    int grid0 = 0;               // workgroups of level 0's restriction (0: the solver's)
    float omega = 0.7f;          // smoother damping
    float gamma = 1.5f;          // coarse-correction scale
};
// One V-cycle on the solver's residual r0 (solve order of level 0): on entry
// z0 holds M r0 (the update kernel's block product, the pre-smoother before
// damping); on exit z0 holds the cycle's result and the solver's kSlotRz
// partials (L0) hold r0^T z0.  st: the state slot the update kernel read.
void launch_vcycle(const MgPlan& P, const Launch& L0, const SysState* st, const float* r0, float* z0);
// z = M r (the level-0 block product alone).
void launch_mg_block_apply(const Launch& L, const BlockPre* M, const float* r, float* z);
// Plain aggregation of a natural-order CSR (n rows): d = |diagonal|,
// strong[e] = entry e is a strong coupling (theta), then rounds of root
// selection (state: 0 undecided, 1 root, 2 excluded; m1, near: work;
// *left = rows still undecided after the round).
void launch_mg_absdiag(int n, const int* rp, const int* ci, const double* va, double* d, hipStream_t s);
void launch_mg_strength(int n, const int* rp, const int* ci, const double* va, const double* d, double theta,
                        unsigned char* strong, hipStream_t s);
void launch_mg_mis_round(int n, const int* rp, const int* ci, const unsigned char* strong, int salt,
                         unsigned char* state, unsigned long long* m1, unsigned char* near, int* left, hipStream_t s);
// flag[i] = row i is a root.
void launch_mg_flag_roots(int n, const unsigned char* state, int* flag, hipStream_t s);
// agg1 / agg: the two joins (rid: the roots' ranks); flag[i] = row i is
// left over.
void launch_mg_join(int n, const int* rp, const int* ci, const double* va, const unsigned char* strong,
                    const unsigned char* state, const int* rid, int* agg1, int* agg, int* flag, hipStream_t s);
// Left-over rows: agg = *nroots + pos.
void launch_mg_left(int n, const int* nroots, const int* pos, int* agg, hipStream_t s);
// Member lists (natural order): cnt = aggregate sizes; after the scan to
// mptr, mem = each aggregate's rows ascending (cur: work, nagg entries).
void launch_mg_member_count(int n, const int* agg, int nagg, int* cnt, hipStream_t s);
void launch_mg_member_fill(int n, const int* agg, int nagg, const int* mptr, int* cur, int* mem, hipStream_t s);
// The same lists per coarse solve row p (perm: coarse solve -> natural, or
// nullptr), as solve positions of the finer level (inv_fine: natural ->
// solve): cnt first, then after the scan to mp, mpos.
void launch_mg_member_solve_count(int np, const int* perm, const int* mptr, int* cnt, hipStream_t s);
void launch_mg_member_solve_pos(int np, const int* perm, const int* mptr, const int* mem, const int* inv_fine,
                                const int* mp, int* mpos, hipStream_t s);
// Smoother blocks cut from aggregates: cnt = pieces per aggregate
// (ceil(size / 8)); after the scan to poff, slots = the pieces (consecutive
// members, sizes as equal as possible), units x 8 rows, -1 padding.
void launch_mg_piece_count(int nagg, const int* mptr, int* cnt, hipStream_t s);
void launch_mg_piece_fill(int nagg, const int* mptr, const int* mem, const int* poff, int* slots, hipStream_t s);
// Galerkin product P^T A P over the member lists (units coarse rows).
void launch_mg_galerkin_count(int units, const int* mptr, const int* mem, const int* rp, int* cnt, hipStream_t s);
void launch_mg_galerkin_fill(int units, const int* mptr, const int* mem, const int* rp, const int* ci,
                             const double* va, const int* agg, const int* off, int* ccol, double* cval, int* dcnt,
                             hipStream_t s);
void launch_mg_galerkin_compact(int units, const int* off, const int* crp, const int* ccol, const double* cval,
                                int* cci, double* cva, hipStream_t s);
void launch_mg_inv_perm(int n_pad, const int* perm, int* inv, hipStream_t s);
// Exact inverses of the diagonal 8x8 blocks (solve order) as full FP32 blocks.
void launch_block_exact_full(const SellView& A, int nq, float* full, hipStream_t s);
// out = inverse of the n x n SPD matrix of a CSR (FP64 Gauss-Jordan, one
// kernel per block of 8 pivots between two buffers, rounded to FP32); work:
// 2 n * n doubles; *bad = 1 on a non-positive pivot (bad must be zero on
// entry).
void launch_mg_dense_inverse(int n, const int* rp, const int* ci, const double* va, double* work, float* out,
                             int* bad, hipStream_t s);

// ---- plain SpMV (no reduction) ----
// Streaming reference: read `bytes` of src once (dst null) or copy them to
// dst, grid workgroups, between events e0 / e1 on stream s.  nontemporal:
// reads 1 = non-temporal loads; copies 1 = non-temporal stores, 2 =
// non-temporal loads and stores.
void launch_stream(int grid, const void* src, void* dst, size_t bytes, unsigned* sink, hipStream_t s,
                   hipEvent_t e0, hipEvent_t e1, int nontemporal = 0);
void launch_spmv_f32(const Launch& L, const SellView& A, const float* x, float* y);
// FP64 SpMV with exact FP64 values (val64) or hi+lo pairs (val + val_lo).
void launch_spmv_f64(const Launch& L, const SellView& A, bool use_lo, const double* x, double* y);
// r = b - A x in FP64 (hi+lo values if use_lo, else exact val64),
// out_rr[sys] = r^T r.  Also writes the FP32 copy of r into r32 (may be null).
void launch_residual_f64(const Launch& L, const SellView& A, bool use_lo, const double* x,
                         const double* b, double* r, float* r32, double* out_rr);
// Residual replacement: r = fp32(b - A x) (FP64 SpMV), z = M r, and the rz /
// rr partials of the next scalar step rewritten from them;
// out[3 sys + 0..2] = (rz, rr, FP64 true r^T r) of each system.
void launch_pcg_replace(const Launch& L, const SellView& A, bool use_lo, const double* x, const double* b,
                        const float* minv, const BlockPre* blockM, float* r, float* z, double* out,
                        const MgPlan* mg = nullptr);
// x64 += (double) d; d = 0.
void launch_fold_f32_into_f64(int n_pad, float* d, double* x, hipStream_t s);
// x64 += (double) d32 over the solve-order rows.
void launch_axpy_f64_f32(int n_pad, const float* d, double* x, hipStream_t s);

// ---- natural-order FP64 element kernels behind the drop-in C-ABI ----
// Each writes per-workgroup partials (partials: >= 2 * grid doubles) and a
// one-workgroup kernel sums them in workgroup order into out[0..1].
void launch_init_pcg_f64(int n, const double* b, const double* ax, const double* m_inv,
                         double* r, double* z, double* p, double* out, double* partials, int grid,
                         hipStream_t s);
void launch_dot_f64(int n, const double* a, const double* b, double* out, double* partials, int grid,
                    hipStream_t s);
void launch_update_xrz_f64(int n, double* x, const double* p, double* r, const double* ap,
                           const double* m_inv, double* z, double alpha, double* out,
                           double* partials, int grid, hipStream_t s);
void launch_update_p_f64(int n, const double* z, double* p, double beta, hipStream_t s);

// ---- permutation between natural and solve order ----
// dst[i] = src[perm[i]] (0 where perm[i] < 0), i < n_pad, with conversion.
void launch_gather_d2f(int n_pad, const int* perm, const double* src, float* dst, hipStream_t s);
void launch_gather_d2d(int n_pad, const int* perm, const double* src, double* dst, hipStream_t s);
void launch_gather_f2f(int n_pad, const int* perm, const float* src, float* dst, hipStream_t s);
// dst[perm[i]] = src[i] for perm[i] >= 0.
void launch_scatter_f2d(int n_pad, const int* perm, const float* src, double* dst, hipStream_t s);
void launch_scatter_d2d(int n_pad, const int* perm, const double* src, double* dst, hipStream_t s);
void launch_scatter_f2f(int n_pad, const int* perm, const float* src, float* dst, hipStream_t s);
void launch_convert_f2d(int n, const float* src, double* dst, hipStream_t s);
// Fills a float or int buffer (graph-safe replacement for a memset).
void launch_fill_f32(size_t n, float v, float* dst, hipStream_t s);
void launch_fill_u32(size_t n, unsigned v, unsigned* dst, hipStream_t s);

// ---- learned block preconditioner ----
// Block features of the solve-order 8-row blocks of a block_rows = 8 matrix:
// feat [n_pad/8 x 48] (upper S_bb, off-block |s| sums, 4 zeros), scale
// [n_pad] = D^-1/2 per solve row; dis [n_pad] is FP64 scratch.
void launch_block_features(const SellView& A, double* dis, float* feat, float* scale, hipStream_t s);
// Full FP32 solve-order blocks (64 floats each) -> factored BlockPre.
void launch_block_factor(int nblocks, const float* full, BlockPre* out, hipStream_t s);
// Identity rows/columns for the padding rows (perm < 0) of solve-order 8x8 blocks.
void launch_mask_padding_blocks(int n_pad, const int* perm, float* blocks, hipStream_t s);

// ---- Jacobi iteration (state machine in JacState, two slots) ----
// x_new = D^-1 (b - R x) on the solve-order vectors.  Launch j of a run
// passes x_old = buffer j % 2 and x_new = buffer (j + 1) % 2; update j runs
// in launch j, so a system stopped after k updates holds its x in buffer
// k % 2.  Each workgroup leaves its max |x_new - x_old| in partial slot
// slot_write and every workgroup reads the previous update's from slot_read
// (launch j: write j % 2, read (j + 1) % 2): with a single slot, a
// workgroup that finished early could overwrite a maximum that a workgroup
// starting late had still to read for the stop test.
void launch_jacobi_iter(const Launch& L, const SellView& A, const JacState* st_cur, JacState* st_next,
                        const float* b, const float* dinv, const float* x_old, float* x_new, int slot_read,
                        int slot_write);

// The fixed-count form: one update with no stop test (partial maxima of
// |x_new - x_old| in slot 0), and the state after `iters` such updates.
void launch_jacobi_sweep(const Launch& L, const SellView& A, const float* b, const float* dinv,
                         const float* x_old, float* x_new);
void launch_jacobi_finish(const Launch& L, int nsys, JacState* st, int iters);

// ---- stream-ordered FP32 device operations (solve-order vectors) ----
// metrics: [2 sys] = r^T z, [2 sys + 1] = r^T r; out: [sys] = a^T b.
void launch_op_init(const Launch& L, int nsys, const float* b, const float* ax, const float* minv, float* r,
                    float* z, float* p, double* metrics);
void launch_op_dot(const Launch& L, int nsys, const float* a, const float* b, double* out);
void launch_op_update_xrz(const Launch& L, int nsys, float* x, const float* p, float* r, const float* ap,
                          const float* minv, float* z, double* metrics, double alpha);
void launch_op_update_p(const Launch& L, const float* z, float* p, double beta);

}  // namespace xp

/*
 * xplus_capi.h -- C-ABI of the MI355X Jacobi-PCG library (libxplus_amd.so).
 *
 * Plain C types only: pointers, sizes, ints, doubles.  No HIP or torch types
 * cross this boundary (a stream is passed as `void*` holding a hipStream_t,
 * NULL = the library's own stream).
 *
 * Three layers:
 *   1. The reference's kernel call surface, same names and signatures.
 *   2. Device-resident matrices and stream-ordered operations (xp_*).
 *   3. The solver: graph-captured Jacobi-PCG, batched systems, FP64
 *      refinement, status and trace readback.
 * Plus host helpers (dataset loader, generators) used by the command-line programs.
 *
 * Unless stated otherwise an xp_* function returns 0 on success and a
 * negative code on failure; xp_last_error() then returns the message.
 */
#ifndef XPLUS_CAPI_H
#define XPLUS_CAPI_H

#ifdef __cplusplus
extern "C" {
#endif

/* ===================================================================== *
 * 1. Drop-in kernel call surface.
 *
 * Replaces archived/include/cg_kernels.hpp:9-54 (bodies
 * archived/kernels/cg_kernels.cpp:43-158): host `double*` buffers owned by
 * the caller, synchronous (each call has finished when it returns), no
 * return code, a silent no-op for n < 0.  Differences, by design:
 *   - no n <= 1024 (kMaxN) limit (archived/kernels/cg_kernels.cpp:9-12);
 *   - the work runs on the GPU in FP64 (exact FP64 values, FP64 vectors),
 *     reductions in a fixed parallel order instead of a sequential loop;
 *   - a malformed CSR (which the reference would read out of bounds) makes
 *     the call a no-op and records the reason for xp_last_error().
 * The SpMV caches the device copy of the last matrix seen and reuses it
 * while a call's three arrays hold the same bytes (compared with a host
 * copy, so in-place edits are seen).
 * ===================================================================== */

/* y = A x, CSR.  archived/include/cg_kernels.hpp:11-16 */
void spmv_csr_kernel(const int* row_ptr, const int* col_idx, const double* values,
                     const double* x, double* y, int n);
/* Same contract; the reference's local version is CSR as well
 * (archived/kernels/cg_kernels.cpp:58-71).  cg_kernels.hpp:18-23 */
void spmv_blocked_kernel(const int* row_ptr, const int* col_idx, const double* values,
                         const double* x, double* y, int n);
/* r = b - ax; z = m_inv r; p = z; metrics = {r^T z, r^T r}.  cg_kernels.hpp:25-32 */
void init_pcg_kernel(const double* b, const double* ax, const double* m_inv, double* r,
                     double* z, double* p, double* metrics, int n);
/* out[0] = a^T b.  cg_kernels.hpp:34-37 */
void dot_kernel(const double* a, const double* b, double* out, int n);
/* x += alpha p; r -= alpha ap; z = m_inv r; metrics = {r^T z, r^T r}.  cg_kernels.hpp:39-47 */
void update_xrz_kernel(double* x, const double* p, double* r, const double* ap,
                       const double* m_inv, double* z, double* metrics, double alpha, int n);
/* p = z + beta p.  cg_kernels.hpp:49-52 */
void update_p_kernel(const double* z, double* p, double beta, int n);

/* Drops the cached device matrix of the drop-in SpMV (also done
 * automatically when the arrays' bytes change). */
void xp_legacy_reset(void);

/* ===================================================================== *
 * Errors and device.
 * ===================================================================== */
/* Copies the calling thread's last error message; returns its length. */
int xp_last_error(char* buf, int len);
/* Number of visible GPUs (0 without one). */
int xp_device_count(void);

/* ===================================================================== *
 * 2. Device-resident SELL-C-sigma matrices.
 *
 * The matrix is converted once (host) and kept on the device in "solve
 * order": rows sorted by length inside windows of `sigma` rows, applied as
 * a symmetric permutation and padded to n_pad rows.  Vectors passed to the
 * stream-ordered operations below are device pointers in solve order of
 * length n_pad; xp_to_solve_order / xp_from_solve_order convert.
 * ===================================================================== */
typedef struct xp_matrix xp_matrix;

typedef struct {
    int sigma;   /* sort window in rows, multiple of 64; 0 = library default */
    int keep_lo; /* keep the FP32 low-part stream (FP64-accurate SpMV, refinement) */
    int keep64;  /* keep exact FP64 values (drop-in FP64 SpMV) */
    int grid;    /* workgroups of the streaming kernels; 0 = library default */
    int spmv_variant; /* SpMV inner loop: 0 rolled loop, 1 fixed-width unrolled,
                         2 unrolled with non-temporal loads of the matrix streams
                         (for matrices that do not stay in the 256 MB MALL) */
    int block_rows;   /* 8: keep blocks of up to 8 rows together as 8 solve rows
                         (block preconditioners); 0/1: rows */
    int blocking;     /* with block_rows = 8: 0 = blocks of 8 consecutive rows;
                         1 = greedy strength aggregates (xp_strength_blocking);
                         2 = strength-ordered aggregates (xp_aggregate_blocking rule 2);
                         3 = parallel strength-ordered aggregates, computed on the
                             device (implies device_build);
                         4 = the same with seeds that precede only their
                             distance-1 neighbours (fewer rounds on dense graphs) */
    int device_build; /* 1: set the matrix up on the device (xp_matrix_create
                         uploads the CSR; one system only) */
} xp_matrix_opts;

typedef struct {
    int n, n_pad, nslices, nsys, sigma, idx16, grid;
    long long nnz;          /* entries of the CSR input */
    long long stored;       /* stored SELL entries incl. padding */
    long long bytes_spmv;   /* HBM bytes one FP32 SpMV streams (values, columns, x, y) */
} xp_matrix_info;

/* One system (batch of 1).  Host CSR arrays; values are FP64. */
int xp_matrix_create(int n, const int* row_ptr, const int* col_idx, const double* values,
                     const xp_matrix_opts* opts, xp_matrix** out);
/* One system from a CSR already in device memory (device pointers), set up
 * entirely on the device: CSR validation, row facts, blocks (block_rows = 8
 * with blocking 0 or 3) and the SELL arrays, equal array for array to the
 * host builder's for the same blocks.  The matrix keeps its own device copy
 * of the CSR.  sigma <= 4096 (block_rows = 1) or 32768 (block_rows = 8).
 * No reference counterpart: the reference packs its matrices on the host
 * (host/cuper_control_matrix.hpp:130-207); this is the set-up the timed
 * learned solve includes. */
int xp_matrix_create_device(int n, const int* row_ptr, const int* col_idx, const double* values,
                            const xp_matrix_opts* opts, xp_matrix** out);
/* The reference's 4x4 bitmap-block format (include/cg_common.hpp:26-50;
 * packer archived/host/xrt_host.cpp:229-349; SpMV
 * archived/kernels/spmv_blocked_kernel.cpp:19-89 and
 * archived/kernels/pcg_control_kernel.cpp:125-245).  xp_spmv_block has the
 * layout of SpmvBlock: bit 4 * local_row + local_col of the mask
 * indices[2] | indices[3] << 8 marks a stored entry, values compact in
 * ascending bit order, the rest zero.  A layout is (n, window_size,
 * row_ptr, col_idx, blocks): window_size 0 = plain (block-row CSR, row_ptr
 * of ceil(n / 4) + 1 entries: the standalone spmv_blocked_kernel's input),
 * else window-major (the reference uses 512, kSpmvWindowSize) with
 * ceil(n / window_size) * (ceil(n / 4) + 1) pointers, window w's at
 * [w (nbr + 1), (w + 1)(nbr + 1)), each window starting where the last one
 * ends.  Block column bc covers columns 4 bc .. 4 bc + 3; an entry outside
 * the matrix or its window multiplies 0.0. */
typedef struct xp_spmv_block {
    double values[16];
    unsigned char indices[16];
} xp_spmv_block;
/* One system from a bitmap-block layout in device memory (see
 * xp_spmv_block): validated and unpacked to CSR on the device (the CSR of
 * xp_bitmap_unpack, entry for entry), then set up as xp_matrix_create_device
 * sets up that CSR.  nblocks <= 134,217,727.  No reference counterpart: the
 * reference's blocked SpMV reads the layout directly; this puts matrices
 * packed for it under the solver. */
int xp_matrix_create_bitmap(int n, int window_size, const int* row_ptr, const int* col_idx,
                            const xp_spmv_block* blocks, long long nblocks, const xp_matrix_opts* opts,
                            xp_matrix** out);
/* The CSR a one-system matrix was built from (host arrays of n + 1 and
 * nnz entries, nnz from xp_matrix_get_info). */
int xp_matrix_csr_copy(const xp_matrix* m, int* row_ptr, int* col_idx, double* values);
/* stats[5]: device-built (0/1), aggregation rounds, regroup rounds, pieces
 * packed, blocks. */
int xp_matrix_setup_stats(const xp_matrix* m, int* stats);
/* Downloads the SELL arrays (sizes from xp_matrix_get_info: n_pad, nslices
 * + 1, nslices, then `stored` entries each; any pointer may be NULL, value
 * streams the matrix does not keep are skipped). */
int xp_matrix_sell_copy(const xp_matrix* m, int* perm, int* slice_ptr, int* slice_base, float* val, float* val_lo,
                        short* val_lo16, double* val64, unsigned short* col16, int* col32);
/* A batch of independent systems solved together (block-diagonal).  The
 * arrays are the systems' CSR arrays concatenated: row_ptrs holds
 * sum(ns[k] + 1) entries (each system's own row_ptr, starting at 0),
 * col_idx/values sum(nnz_k) entries with system-local column indices. */
int xp_matrix_create_batch(int nsys, const int* ns, const int* row_ptrs, const int* col_idx,
                           const double* values, const xp_matrix_opts* opts, xp_matrix** out);
void xp_matrix_destroy(xp_matrix* m);
int xp_matrix_get_info(const xp_matrix* m, xp_matrix_info* out);
/* perm[i] = natural (global) row of solve row i, -1 for padding; n_pad entries. */
int xp_matrix_permutation(const xp_matrix* m, int* perm);
/* Changes the workgroup count of the streaming kernels (autotuning). */
int xp_matrix_set_grid(xp_matrix* m, int grid);
int xp_matrix_set_variant(xp_matrix* m, int spmv_variant);

/* Autotuner: times `iters` Jacobi-PCG iterations (the real graph-captured
 * iteration: 3-kernel for fused = 0, fused for 1, the accurate iteration of
 * xp_solve_opts.accurate for 2, which needs keep_lo) for every candidate of SpMV
 * inner loop x workgroup count (x sort window sigma when tune_sigma != 0,
 * which re-sorts the matrix and so must run before any solver is created on
 * it), keeps the fastest and returns it.  Results are cached per matrix
 * signature (n, nnz, content hash) for the life of the process. */
typedef struct {
    int sigma, spmv_variant, grid;
    double us_per_iter;      /* best */
    double us_per_iter_default;
    int candidates;
} xp_tune_report;
int xp_matrix_autotune(xp_matrix* m, int fused, int tune_sigma, int iters, xp_tune_report* out);

/* y = A x, FP32 values and vectors, FP32 accumulation. */
int xp_spmv_f32(xp_matrix* m, const float* x, float* y, void* stream);
/* y = A x in FP64: exact values if keep64, else hi+lo pairs (needs keep_lo). */
int xp_spmv_f64(xp_matrix* m, const double* x, double* y, void* stream);
/* Stream-ordered FP32 device operations on solve-order vectors of n_pad
 * entries: the five fused operations (archived/kernels/cg_kernels.cpp:73-158)
 * on device pointers, for callers that drive their own iteration with
 * xp_spmv_f32.  Each stored value is evaluated in FP64 from the FP32
 * operands (no fused multiply-add) and rounded to FP32 when stored;
 * reductions are FP64 sums of the exact FP32 products in a fixed order, one
 * per system, written to device memory: metrics[2 k] = r^T z,
 * metrics[2 k + 1] = r^T r, out[k] = a^T b.  Padding rows of the inputs
 * must be finite (xp_to_solve_order writes zeros there).  They use the
 * matrix's reduction scratch, as its solvers do: do not overlap them with
 * a solve on the same matrix.
 *   init:       r = b - ax, z = minv r, p = z
 *   update_xrz: x += alpha p, r -= alpha ap, z = minv r
 *   update_p:   p = z + beta p */
int xp_init_pcg_f32(xp_matrix* m, const float* b, const float* ax, const float* minv, float* r, float* z, float* p,
                    double* metrics, void* stream);
int xp_dot_f32(xp_matrix* m, const float* a, const float* b, double* out, void* stream);
int xp_update_xrz_f32(xp_matrix* m, float* x, const float* p, float* r, const float* ap, const float* minv, float* z,
                      double* metrics, double alpha, void* stream);
int xp_update_p_f32(xp_matrix* m, const float* z, float* p, double beta, void* stream);
/* natural (n, FP64) -> solve order (n_pad, FP32), and back. */
int xp_to_solve_order(xp_matrix* m, const double* nat, float* solve, void* stream);
int xp_from_solve_order(xp_matrix* m, const float* solve, double* nat, void* stream);
/* Host-pointer convenience: y = A x in FP64 on the GPU for a CSR given once. */
/* The achievable streaming rate of the GPU, for roofline figures: a
 * project-written kernel reads (mode 0), copies (mode 1) or reads with
 * non-temporal loads, four in flight per thread (mode 2), copies with
 * non-temporal stores (mode 3) or non-temporal loads and stores (mode 4),
 * four in flight per thread, `bytes` with
 * 16-byte loads in `grid` workgroups; average microseconds per launch over
 * `reps` launches, each between its own events.  No reference counterpart. */
int xp_stream_time(unsigned long long bytes, int mode, int grid, int reps, double* avg_us);
int xp_csr_spmv_f64(int n, const int* row_ptr, const int* col_idx, const double* values,
                    const double* x, double* y);

/* y = A x straight from a bitmap-block layout in device memory (all
 * pointers device, y n doubles; stream-ordered).  FP64, each row summed in
 * the format's order from +0.0 with product and sum rounded separately:
 * bit for bit the reference's sequential loops.  Replaces the HLS
 * spmv_blocked_kernel(b_row_ptr, b_col_idx, blocks, x, y, n)
 * (archived/kernels/spmv_blocked_kernel.cpp:131-161; window_size 0) and the
 * windowed SpMV of pcg_control_kernel (spmv_blocked_windowed,
 * archived/kernels/pcg_control_kernel.cpp:317-506; window_size 512).  The
 * layout is not validated here (xp_matrix_create_bitmap and
 * xp_bitmap_unpack do): pointers outside [0, nblocks] are clamped and
 * negative block columns read 0.0, so malformed input reads nothing out of
 * bounds. */
int xp_bitmap_spmv_f64(int n, int window_size, const int* row_ptr, const int* col_idx, const xp_spmv_block* blocks,
                       long long nblocks, const double* x, double* y, void* stream);
/* Average microseconds per launch of xp_bitmap_spmv_f64 over `reps`
 * launches after `warmup`, each between its own HIP events on the library
 * stream (device pointers, x and y of n doubles). */
int xp_bitmap_spmv_time(int n, int window_size, const int* row_ptr, const int* col_idx, const xp_spmv_block* blocks,
                        long long nblocks, const double* x, double* y, int warmup, int reps, double* avg_us);

/* Device memory and timing helpers for drivers that hold no other GPU
 * runtime (tests, benchmarks): plain hipMalloc/hipMemcpy on the library's
 * stream. */
int xp_device_alloc(unsigned long long bytes, void** out);
int xp_device_free(void* ptr);
/* The library's own device buffers (matrices, solvers, hierarchies) come
 * from a caching allocator (no reference counterpart): released blocks of
 * at least 256 KB are kept and reused, after one device synchronisation,
 * instead of going back to hipFree (a hipMalloc / hipFree pair of a few MB
 * costs 0.2-0.4 ms).  xp_device_pool_trim returns every cached block to the
 * runtime; xp_device_pool_bytes reports how many bytes are cached. */
int xp_device_pool_trim(void);
int xp_device_pool_bytes(unsigned long long* bytes);
int xp_memcpy_htod(void* dst, const void* src, unsigned long long bytes);
int xp_memcpy_dtoh(void* dst, const void* src, unsigned long long bytes);
/* Device-to-device copy (the runtime's copy kernel): a transfer of known
 * size, e.g. to calibrate memory-traffic counters. */
int xp_memcpy_dtod(void* dst, const void* src, unsigned long long bytes);
int xp_device_synchronize(void);
/* Times `reps` back-to-back launches of the FP32 SpMV (kernel k_spmv_f32)
 * with HIP events recorded on the stream the kernels are launched on (the
 * library stream), after `warmup` untimed launches; x is filled with a
 * fixed pattern.  avg_us receives the mean time per launch. */
int xp_spmv_f32_time(xp_matrix* m, int warmup, int reps, double* avg_us);
/* The same with cold caches: a device-to-device copy of flush_bytes before
 * every timed launch evicts the L2s and the MALL, so the matrix and x are
 * read from HBM (the warm figure above can be served partly from the MALL). */
int xp_spmv_f32_time_cold(xp_matrix* m, int reps, unsigned long long flush_bytes, double* avg_us);

/* ===================================================================== *
 * 3. Solver.
 *
 * Jacobi-PCG with the semantics of the reference's on-chip loop
 * (kernels/cuper_pcg_control_kernel.cpp:1069-1234) and golden
 * (src/CgSolverGolden.hpp:42-124): stop when r^T r <= tau on the
 * recursively updated residual; breakdown on non-finite or |.| <= 1e-30
 * pAp / rz and non-finite alpha / beta.  FP32 vectors and SpMV, FP64
 * reductions and scalars.  One iteration = 3 kernels, captured in a HIP
 * graph of `graph_iters` iterations; the host checks the state once per
 * graph launch.
 * ===================================================================== */
typedef struct xp_pcg xp_pcg;

enum {
    XP_CONVERGED = 0,
    XP_MAX_ITER = 1,
    XP_BREAKDOWN = 2,
    XP_STAGNATION = 3,
    XP_RUNNING = -1
};

typedef struct {
    double tau;        /* absolute bound on r^T r; used when rel_tol <= 0 */
    double rel_tol;    /* if > 0: tau_k = (rel_tol * ||b_k||)^2 per system */
    int max_iters;     /* 0 -> max(4 n_k, 1000) (src/CgSolverGolden.hpp:47) */
    int graph_iters;   /* iterations per graph launch; 0 = 64 */
    int use_graph;     /* 0: launch kernels directly, graph_iters per host check (short solves: no graph
                          to capture and instantiate) */
    int trace_cap;     /* trace rows kept per system */
    int stall_limit;   /* 0 = off; else stop with XP_STAGNATION after this many
                          iterations without halving the best r^T r */
    int fused;         /* 1: two kernels per iteration (p update folded into the
                          SpMV via A p = A z + beta A p_old); 0: three kernels */
    int accurate;      /* 1: FP64-accumulated SpMV (values: FP32 hi + a 16-bit
                          remainder, ~39 bits; exact FP64 with keep64), FP64 p,
                          A p and x, FP32 r and z; two kernels per iteration.
                          The recursive residual then tracks b - A x, so
                          ill-conditioned systems reach a true 1e-6 without
                          refinement.  2: the same with the FP32 hi + lo pair
                          (~48 bits), which xp_pcg_run_reliable also uses for
                          targets below 1e-8.  Needs keep_lo or keep64.  0:
                          FP32 iteration.  True residuals always use hi + lo. */
    int full_work_map; /* batches: 0 (default) marks the workgroups of systems
                          that have stopped idle between graph launches (they
                          return at once instead of streaming a first slice and
                          re-summing partials); 1 keeps every workgroup */
} xp_solve_opts;

typedef struct {
    int status;        /* XP_* */
    int iterations;
    double rz, rr, pap, alpha, beta;
    double rr0;        /* r^T r after init */
    double tau;
    int fallback;      /* 1: the block preconditioner stagnated or broke down and
                          the solve was redone with plain Jacobi */
} xp_status;

/* dot_kernel on device vectors, allocating nothing: out[0] = a^T b (FP64,
 * the same kernels, split and summation order as dot_kernel, so the same
 * bits), stream-ordered on `stream` (NULL: the library stream).  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Its per-workgroup partials live in one fixed
 * 8 KiB device array per GPU (at most 1024 partials; larger splits are
 * refused with an error), held until the process ends; calls that overlap
 * on one GPU share that array and race.  n < 0: no-op (as dot_kernel). */
int xp_dot_f64_device(const double* a, const double* b, double* out, int n, void* stream);

int xp_pcg_create(xp_matrix* m, const xp_solve_opts* opts, xp_pcg** out);
void xp_pcg_destroy(xp_pcg* s);
/* Uploads b and x0 (host, natural order, FP64, sum of n_k entries; x0 may
 * be NULL for zero) and derives tau per system.  Not part of a timed solve. */
int xp_pcg_set_rhs(xp_pcg* s, const double* b, const double* x0);
/* The same from device memory (natural order, FP64; x0 may be NULL). */
int xp_pcg_set_rhs_device(xp_pcg* s, const double* b, const double* x0);
/* Runs init + iterations until every system stops.  Device-resident. */
int xp_pcg_run(xp_pcg* s);
/* Runs init + exactly `iters` iterations (graph-captured as one launch),
 * stopping early only where the state machine stops a system. */
int xp_pcg_run_fixed(xp_pcg* s, int iters);
/* Mixed-precision refinement: outer FP64 residual b - A x (hi+lo values)
 * and FP64 solution update on the GPU, inner FP32 PCG to
 * inner_rel * ||r_outer||, until ||b - A x|| <= outer_rel * ||b|| per system.
 * Needs keep_lo.  Writes the outer true relative residuals to rel_res. */
int xp_pcg_run_refined(xp_pcg* s, double outer_rel, double inner_rel, int max_outer,
                       double* rel_res, int* outer_steps, int* total_iters);
/* FP32 PCG with residual replacement (reliable updates) and an FP64
 * solution: at graph-launch boundaries, when a system's r^T r has fallen by
 * delta^2 since its last replacement (or it converged on the recursive
 * residual), x64 += x32 and r, z are replaced by the FP64 true residual of
 * x64 (hi+lo values) while p is kept, so the Krylov recurrence is not
 * restarted.  A system stops when ||b - A x64|| <= rel_tol * ||b||.  Needs
 * keep_lo; runs the fused iteration form; delta in [0, 1): 0.1 suits the
 * FP32 iteration; 0 replaces only when a system has converged on its
 * recursive residual, which suits the accurate iteration (its recursive
 * residual tracks the true one, and each replacement perturbs the Krylov
 * recurrence).
 * With a block preconditioner and stall_limit > 0 the true residual judges
 * stagnation: every stall_limit iterations (at the next graph boundary) the
 * FP64 ||b - A x|| of every active system is computed, and when one has not
 * halved its best r^T r since the previous check, or a system broke down,
 * the solver switches to Jacobi and restarts every unfinished system from
 * its current x (fallback = 1 in xp_status; the blocks are used again by
 * the next run).  The recursive stagnation watch is off in this mode.
 * Writes true relative residuals, the number of replacements and the
 * largest iteration count. */
int xp_pcg_run_reliable(xp_pcg* s, double rel_tol, double delta, double* rel_res, int* replacements,
                        int* total_iters);
int xp_pcg_get_status(xp_pcg* s, int sys, xp_status* out);
/* Solution in natural order, FP64 (sum of n_k entries). */
int xp_pcg_get_x(xp_pcg* s, double* x);
/* The solution into device memory (natural order, FP64, n entries). */
int xp_pcg_get_x_device(xp_pcg* s, double* x);
/* Trace rows {alpha, beta, rz, rr} of system `sys`; returns rows copied. */
int xp_pcg_get_trace(xp_pcg* s, int sys, double* rows, int cap);
/* FP64 true residual ||b - A x||_2 per system on the GPU. */
int xp_pcg_true_residual(xp_pcg* s, double* out_norms);
/* Preconditioner.  Jacobi (z = r / diag) by default.  A block preconditioner
 * needs a matrix built with block_rows = 8; blocks are the aligned groups of
 * 8 natural rows of each system (the last one may be partial), each an 8x8
 * row-major FP32 matrix M_b applied as z_b = M_b r_b.
 *   mode 0: Jacobi;  mode 1: exact block Jacobi (M_b = A_bb^-1, FP64 Cholesky
 *   on the host);  mode 2: user SPD blocks from `blocks` (nblocks x 64
 *   floats, block order).  The device stores each block factored,
 *   M_b = S C C^T S with s_i = sqrt(M_b,ii) and C the Cholesky factor of
 *   S^-1 M_b S^-1 in FP16 (14 bytes per row instead of 32); a block that is
 *   not positive definite keeps only its diagonal.  With a block
 *   preconditioner and stall_limit > 0, a stagnating or breaking-down solve
 *   is redone with Jacobi (fallback). */
int xp_pcg_set_preconditioner(xp_pcg* s, int mode, const float* blocks, long long count);
/* count: the number of floats in `blocks` (mode 2: must be
 * xp_matrix_num_blocks(m) * 64, else the call fails; ignored otherwise). */
/* The operator the device applies for each block, evaluated in FP64 from
 * the stored FP16 factor and FP32 scales (S C C^T S), in block order:
 * out[64 b + 8 i + j], xp_matrix_num_blocks(m) * 64 doubles; padding slots
 * hold identity rows and columns.  Fails when no block preconditioner is
 * installed.  (No reference counterpart: a check of the learned
 * preconditioner against the FP64 block-PCG oracle.) */
int xp_pcg_block_operator(xp_pcg* s, double* out);
/* Number of 8-row blocks of the matrix (all systems): sum ceil(n_k / 8)
 * for consecutive-row blocks, the aggregate count for strength blocking.
 * Block-ordered arrays (mode 2 blocks, xp_matrix_block_features) list the
 * blocks of system 0, then system 1, ... */
int xp_matrix_num_blocks(const xp_matrix* m);
/* The rows of every block: rows[8 b + j] = global natural row of slot j of
 * block b, -1 for a padding slot (real rows first). */
int xp_matrix_block_rows(const xp_matrix* m, int* rows);
/* Greedy strength aggregation of one CSR system into blocks of at most 8
 * rows (no GPU): s_ij = |a_ij| / sqrt(a_ii a_jj); seeds in row order, a
 * block grows by the unassigned row of largest summed strength to it (ties:
 * smaller index) up to 8 rows; rows of partial blocks are regrouped among
 * themselves the same way, the remaining partial blocks packed first-fit;
 * rows ascending inside a block, blocks ordered by their first row.
 * slots: capacity 8 n; writes 8 * nblocks entries (-1 = padding). */
int xp_strength_blocking(int n, const int* row_ptr, const int* col_idx, const double* values, int* slots,
                         int* nblocks);
/* The same with a choice of rule: 1 = xp_strength_blocking; 2 = seeds in
 * decreasing order of the row's total off-diagonal strength t_i (ties: the
 * smaller row), a block grows by the unassigned row with the largest
 * (strength to the block) / t_j (ties: the smaller index); the regrouping
 * and packing steps as rule 1. */
int xp_aggregate_blocking(int n, const int* row_ptr, const int* col_idx, const double* values, int rule, int* slots,
                          int* nblocks);
/* Average device time (us) of each kernel of one iteration, over `iters`
 * directly launched iterations from x0 with the convergence test off, HIP
 * events recorded on the solver's stream between kernels: us[0] the SpMV
 * kernel (fused or not), us[1] update_xrz, us[2] update_p (three-kernel
 * form only).  Leaves the solver's vectors as after those iterations. */
int xp_pcg_kernel_times(xp_pcg* s, int iters, double* us, int* nkernels);
/* The same measurement split by iteration parity: us6[0..2] the kernels'
 * averages over the even iterations, us6[3..5] over the odd ones (the
 * accurate iteration's update kernel adds x only in odd iterations). */
int xp_pcg_kernel_times_split(xp_pcg* s, int iters, double* us6, int* nkernels);
/* Device time of the last xp_pcg_run* (ms, HIP events) and launches made. */
int xp_pcg_last_timing(xp_pcg* s, double* ms, int* graph_launches);

/* ===================================================================== *
 * Learned 8x8 block-Jacobi preconditioner (new: the reference has none).
 *
 * In the Jacobi-scaled matrix S = D^-1/2 A D^-1/2, per aligned 8-row block b
 * of a system: Ahat = S_bb; features (48) = upper triangle of Ahat (36) +
 * off-block |s| row sums (8) + 4 zeros; MLP 48 -> 64 -> 64 -> 48 (ReLU,
 * ReLU, linear), outputs 0..35 = lower triangle of L; block
 * M_b = D^-1/2 (L L^T + 1e-3 I) D^-1/2 (SPD by construction; L = I is plain
 * Jacobi, the initialisation).  Forward and input-gradient GEMMs are BF16
 * MFMA with FP32 accumulation; weight gradients FP32; Adam.  Training loss
 * (block SPAI): mean over blocks of ||Mh Ahat - I||_F^2 + 0.5 ||Mh S_off||_F^2.
 * Parameters are one flat FP32 vector: W1[48][64], b1[64], W2[64][64],
 * b2[64], W3[64][48], b3[48] (row-major, input index first).
 * ===================================================================== */
typedef struct xp_mlp xp_mlp;
typedef struct xp_blockset xp_blockset;
int xp_mlp_create(unsigned long long seed, xp_mlp** out);
void xp_mlp_destroy(xp_mlp* net);
int xp_mlp_num_params(void);
int xp_mlp_get_params(xp_mlp* net, float* out);
int xp_mlp_set_params(xp_mlp* net, const float* params);
/* Blocks of a host CSR system (no GPU needed); append concatenates sets. */
int xp_blockset_from_csr(int n, const int* row_ptr, const int* col_idx, const double* values, xp_blockset** out);
/* The same over the blocks of a blocking (0: consecutive rows, 1 / 2: strength aggregates, rule 1 / 2),
 * with the off-block weight of the training loss (< 0: 0.5). */
int xp_blockset_from_csr_blocked(int n, const int* row_ptr, const int* col_idx, const double* values, int blocking,
                                 double lambda, xp_blockset** out);
/* The same over given blocks (units x 8 rows of slots, -1 padding; e.g.
 * xp_matrix_block_rows of a device-aggregated matrix, blocking 3). */
int xp_blockset_from_csr_slots(int n, const int* row_ptr, const int* col_idx, const double* values, const int* slots,
                               int units, double lambda, xp_blockset** out);
int xp_blockset_append(xp_blockset* dst, const xp_blockset* src);
int xp_blockset_size(const xp_blockset* set);
/* feat: size x 48, ahat: size x 64, hmat (Ahat^2 + 0.5 G): size x 64, scale (D^-1/2): size x 8. */
int xp_blockset_copy(const xp_blockset* set, float* feat, float* ahat, float* hmat, float* scale);
void xp_blockset_free(xp_blockset* set);
/* Mean loss over samples [first, first+count) and its gradient (flat). */
int xp_mlp_loss_grad(xp_mlp* net, const xp_blockset* set, int first, int count, double* loss, float* grad);
/* Network outputs (count x 48) and blocks M_b (size x 64). */
int xp_mlp_forward(xp_mlp* net, const xp_blockset* set, int first, int count, float* out);
int xp_mlp_blocks(xp_mlp* net, const xp_blockset* set, float* blocks);
/* `steps` Adam steps on contiguous batches of `batch` samples; loss per step. */
int xp_mlp_train(xp_mlp* net, const xp_blockset* set, int steps, int batch, float lr, unsigned long long seed,
                 double* loss_history);
/* Builds the blocks of the solver's matrix with the network and installs
 * them as its preconditioner (like xp_pcg_set_preconditioner mode 2).  All
 * on the device: features from the resident SELL matrix, BF16 MFMA
 * inference, blocks written in solve order; padding rows stay identity.
 * The matrix must be built with block_rows = 8. */
int xp_pcg_set_mlp_preconditioner(xp_pcg* s, xp_mlp* net);

/* Multilevel aggregation preconditioner (no reference counterpart: it
 * extends the block-Jacobi preconditioner the reference's Jacobi-PCG loop
 * would take in place of its inverse diagonal, src/CgSolverGolden.hpp:33-40).
 * The installed 8x8 blocks (exact or learned) become the smoother of level
 * 0.  Each level's coarse space is a plain aggregation of its rows: on the
 * strength graph (a_ij^2 >= theta^2 |a_ii| |a_jj|), roots form a maximal
 * distance-2 independent set chosen in parallel rounds, a root's strong
 * neighbours join it, every other row joins its strongest joined neighbour;
 * level l + 1 is the Galerkin product P^T A P over those aggregates (P
 * piecewise constant), set up on the device with its own 8-row smoother
 * blocks (coarse_blocking 3 or 4: the parallel strength aggregation of
 * xp_matrix_opts.blocking; 1: pieces of the level's own aggregates; levels
 * of at most consecutive_rows rows take 8 consecutive rows) and exact block
 * factors, down to a dense coarsest level
 * of at most coarse_max rows (inverted once).  Each PCG step then applies
 * one symmetric V-cycle: z = omega M r, coarse correction scaled by gamma,
 * one post-smoothing step; FP32 vectors inside the cycle.  Needs the
 * accurate iteration on one system set up on the device with block_rows =
 * 8, and a block preconditioner already installed.  A fallback to Jacobi
 * turns it off with the blocks. */
typedef struct {
    int coarse_max;       /* rows of the dense coarsest level, at most (8..4096); <= 0 removes the hierarchy */
    double omega;         /* smoother damping, 0 < omega < 2 */
    double gamma;         /* coarse-correction scale, 0 < gamma < 4 */
    int coarse_blocking;  /* smoother blocks of the coarse levels: 3 or 4 (see xp_matrix_opts.blocking), or 1:
                             pieces of the level's own aggregates (ceil(size / 8) pieces of consecutive rows each,
                             sizes as equal as possible; no aggregation rounds of their own) */
    int consecutive_rows; /* coarse levels of at most this many rows take 8 consecutive rows per block */
    double theta;         /* strength threshold of the aggregation, 0 <= theta < 1 */
} xp_multilevel_opts;
int xp_pcg_set_multilevel(xp_pcg* s, const xp_multilevel_opts* opts);
/* Levels of the hierarchy (*nlev, 0 when none): rows and stored entries of
 * the first `cap` levels (level 0 is the solver's matrix, the last the dense
 * coarsest one) and the set-up time in ms. */
int xp_pcg_multilevel_info(const xp_pcg* s, int cap, int* nlev, int* n, long long* nnz, double* setup_ms);
/* The set-up's time per level in ms, three entries per level below the
 * coarsest: aggregation (with its member lists), Galerkin product, the
 * coarse level's set-up (device matrix set-up and block factors, or the
 * dense inverse); *count entries. */
int xp_pcg_multilevel_phases(const xp_pcg* s, int cap, double* ms, int* count);
/* Root-selection rounds of each level's aggregation; *count entries. */
int xp_pcg_multilevel_rounds(const xp_pcg* s, int cap, int* rounds, int* count);
/* Level `level` >= 1 as a natural-order CSR (rows = the aggregates of level
 * level - 1) and the aggregate of every natural row of level level - 1;
 * null outputs are skipped (sizes from xp_pcg_multilevel_info). */
int xp_pcg_multilevel_level(const xp_pcg* s, int level, int* row_ptr, int* col_idx, double* values, int* agg);
/* The smoother blocks of SELL level 1 <= level < levels - 1: *units blocks,
 * units x 8 natural rows of the level (-1 padding) into slots (may be
 * null). */
int xp_pcg_multilevel_slots(const xp_pcg* s, int level, int* slots, int* units);
/* z = M r followed by one V-cycle, for device vectors of level 0 (solve
 * order, n_pad entries). */
int xp_pcg_multilevel_apply(xp_pcg* s, const float* r, float* z);
/* Average device microseconds of one V-cycle (with the block product) over
 * `reps` cycles launched back to back on the solver's residual; overwrites
 * the solver's z (the next run recomputes it). */
int xp_pcg_multilevel_time(xp_pcg* s, int reps, double* us);
/* The device-computed features of a block_rows = 8 matrix, in natural block
 * order (xp_matrix_num_blocks of them): feat nblocks x 48, scale nblocks x 8
 * (the same quantities as xp_blockset_copy's feat and scale). */
int xp_matrix_block_features(const xp_matrix* m, float* feat, float* scale);

/* ===================================================================== *
 * Plain Jacobi iteration x_new = D^-1 (b - R x), A = D + R
 * (DLC/Cuper-jacobi-iteration/host/main.cpp:957-1053, the reference's
 * PrepareJacobiVectorsAndR + RunJacobiCpu and its CuperJacobiIteration
 * kernel, kernels/detail/jacobi_controller.hpp:155-283).  FP32 like the
 * reference's VALUE_TYPE: D is the FP32 sum of each row's diagonal entries
 * in CSR order, Diag_inv = 1 / D, R x summed per row in CSR order with one
 * rounding per product and per sum, so every update is the reference's bit
 * for bit.  Without use_tau: exactly max_iters updates with no stop test
 * (the reference's fixed-count contract), then status breakdown if the last
 * update's difference is not finite, else max_iter.  With use_tau the
 * iteration also stops once max_i |x_new - x_old| <= tau (the stop rule of
 * docs/jacobi_iteration.md:99-124), and at the first non-finite difference
 * as breakdown.  Error "missing or zero diagonal at row R".
 * ===================================================================== */
typedef struct xp_jacobi xp_jacobi;
typedef struct xp_jacobi_opts {
    double tau;       /* default 1e-5 (the reference's TAU) */
    int max_iters;    /* default 1 (the reference's MAX_ITERS); 0 allowed */
    int use_tau;      /* 0: fixed count */
    int graph_iters;  /* updates per graph launch (default 32, rounded up to even) */
    int use_graph;
} xp_jacobi_opts;
typedef struct xp_jacobi_status {
    int status;       /* 0 converged, 1 max_iter (kJacobiStatusMaxIter), 2 breakdown */
    int iterations;   /* updates done (Status[2] of the reference kernel) */
    double diff;      /* max |x_new - x_old| of the last update */
} xp_jacobi_status;
int xp_jacobi_create(xp_matrix* m, const xp_jacobi_opts* opts, xp_jacobi** out);
void xp_jacobi_destroy(xp_jacobi* s);
/* b and x0 in natural order (NULL x0: zeros, as the reference starts). */
int xp_jacobi_set_rhs(xp_jacobi* s, const double* b, const double* x0);
int xp_jacobi_run(xp_jacobi* s);
int xp_jacobi_get_status(xp_jacobi* s, int sys, xp_jacobi_status* out);
int xp_jacobi_get_x(xp_jacobi* s, double* x);
int xp_jacobi_last_timing(xp_jacobi* s, double* ms, int* graph_launches);
/* Average device time of one update kernel over `iters` direct launches. */
int xp_jacobi_kernel_time(xp_jacobi* s, int iters, double* us);
/* The reference's right-hand side when a dataset has no b: A * ones in FP32
 * (host/main.cpp:969-973), on the host. */
int xp_rhs_ones_f32(int n, const int* row_ptr, const int* col_idx, const double* values, double* b);

/* ===================================================================== *
 * Host helpers: loader and generators (no GPU needed).
 * ===================================================================== */
typedef struct xp_csr xp_csr;
/* CsrDataset::load (src/CsrDataset.hpp:19-67): same files, same checks,
 * same messages (via xp_last_error). */
int xp_dataset_load(const char* dir, xp_csr** out);
int xp_dataset_save(const xp_csr* a, const char* dir);
/* MatrixMarket ingestion: parse_coordinate_matrix, parse_vector,
 * reference_solution and convert_dataset of
 * scripts/download_suitesparse_data.py:224-349,418-444.  Coordinate
 * real|integer|pattern matrices, general|symmetric; explicit zeros skipped,
 * duplicates summed in file order, symmetric entries mirrored, the leading
 * subset_n x subset_n principal submatrix kept (subset_n <= 0: all), columns
 * ascending, sums equal to 0 dropped.  b is read from b_path (array or
 * coordinate; NULL or "" for none), else b = A x_ref with the converter's
 * x_ref and a sequential FP64 row sum; x0 = 0.  The converter's error
 * messages; a row or column index below 1 is an error. */
typedef struct xp_mtx_info {
    long long nrows, ncols, nnz_reported;  /* the size line */
    long long entries;                     /* entry lines read */
    int field;                             /* 0 real, 1 integer, 2 pattern */
    int symmetry;                          /* 0 general, 1 symmetric */
    int n, m;                              /* shape kept */
    long long nnz;                         /* stored entries */
    int has_diag;                          /* a nonzero diagonal entry is kept */
    double diag_min_abs;                   /* smallest |a_ii| kept (0 when none) */
} xp_mtx_info;
int xp_mtx_load(const char* mtx_path, const char* b_path, long long subset_n, xp_mtx_info* info, xp_csr** out);
/* parse_vector (download_suitesparse_data.py:287-336): writes
 * min(expected_n, subset_n) entries (expected_n when subset_n <= 0). */
int xp_mtx_read_vector(const char* path, int expected_n, long long subset_n, double* out);
/* reference_solution (download_suitesparse_data.py:339-349), n entries. */
int xp_reference_solution(int n, double* x);
/* Binary CSR cache (no counterpart in the reference, which reparses the
 * text files on every run): one file holding row_ptr, col_idx, values, b, x0
 * with a payload checksum.  load_cached reads dir's five text files through
 * the cache at cache_path (NULL or "": dir/csr_cache.xpbin), which is used
 * when the recorded sizes and modification times of the text files match and
 * rewritten otherwise (a read-only location is not an error); *hit says
 * which.  Same validation and messages as xp_dataset_load. */
int xp_dataset_save_binary(const xp_csr* a, const char* path);
int xp_dataset_load_binary(const char* path, xp_csr** out);
int xp_dataset_load_cached(const char* dir, const char* cache_path, int* hit, xp_csr** out);
/* Generators (seeded, deterministic).  b ~ U[-1,1], x0 = 0. */
int xp_gen_diffusion2d(int nx, int ny, double sigma, unsigned long long seed, xp_csr** out);
int xp_gen_diffusion3d(int k, double sigma, unsigned long long seed, xp_csr** out);
int xp_gen_fem2d(int m, double sigma, int n_incl, double contrast, double jitter,
                 unsigned long long seed, xp_csr** out);
int xp_csr_from_arrays(int n, const int* row_ptr, const int* col_idx, const double* values,
                       const double* b, const double* x0, xp_csr** out);
int xp_csr_info(const xp_csr* a, int* n, int* nnz);
int xp_csr_copy(const xp_csr* a, int* row_ptr, int* col_idx, double* values, double* b, double* x0);
void xp_csr_free(xp_csr* a);
/* The reference's Cuper packed matrix format (host/cuper_control_matrix.hpp:
 * 14-207 build_cuper_control_matrix; the same 64-bit word layout as
 * DLC/Cuper/include/Cuper_common.h:535-774): 8192-column batches, rows dealt
 * to 128 PEs, 16 channels of 8 lanes, word = local column << 50 | row code
 * << 32 | FP32 value bits, empty slot 0x3FFFF << 32.  xp_cuper_pack
 * reproduces the reference's packer word for word; xp_cuper_unpack turns
 * any such matrix (16 channel buffers of channel_size words, the batch table
 * sp_element_list_ptr) back into CSR (columns ascending, FP32 values) so it
 * can be built into a device matrix.  At most 16,777,216 rows. */
typedef struct xp_cuper_packed xp_cuper_packed;
int xp_cuper_pack(int n, const int* row_ptr, const int* col_idx, const double* values, xp_cuper_packed** out);
int xp_cuper_packed_info(const xp_cuper_packed* p, int* batch_count, int* matrix_len, int* channel_size);
/* sp_element_list_ptr: batch_count + 1 entries; channels: 16 buffers of
 * channel_size words each (the reference's matrix_fpga_data_[0..15]). */
int xp_cuper_packed_copy(const xp_cuper_packed* p, int* sp_element_list_ptr, unsigned long long* const* channels);
int xp_cuper_packed_create(int batch_count, int channel_size, const int* sp_element_list_ptr,
                           const unsigned long long* const* channels, xp_cuper_packed** out);
int xp_cuper_unpack(const xp_cuper_packed* p, int n, xp_csr** out);
void xp_cuper_packed_free(xp_cuper_packed* p);
/* The bitmap-block format (xp_spmv_block) on the host: the reference's
 * packer and the unpack to CSR. */
typedef struct xp_bitmap xp_bitmap;
/* build_windowed_blocked_spmv_matrix: duplicate CSR entries summed in CSR
 * order, positions whose sum is 0.0 and empty blocks dropped, block
 * columns ascending within a block row.  The host builder's CSR checks. */
int xp_bitmap_pack(int n, const int* row_ptr, const int* col_idx, const double* values, int window_size,
                   xp_bitmap** out);
int xp_bitmap_info(const xp_bitmap* b, int* num_windows, int* num_block_rows, long long* nblocks);
/* row_ptr: num_windows * (num_block_rows + 1) entries; col_idx, blocks: nblocks. */
int xp_bitmap_copy(const xp_bitmap* b, int* row_ptr, int* col_idx, xp_spmv_block* blocks);
void xp_bitmap_free(xp_bitmap* b);
/* CSR of a layout held in host memory: per row the entries y = A x reads,
 * in its order (windows, blocks, bits), entries outside the matrix or their
 * window dropped, stored zeros kept; b = x0 = 0.  Errors "bitmap row_ptr
 * out of range", "bitmap row_ptr must be nondecreasing", "bitmap row_ptr
 * windows must be contiguous", "bitmap col_idx must be non-negative". */
int xp_bitmap_unpack(int n, int window_size, const int* row_ptr, const int* col_idx, const xp_spmv_block* blocks,
                     long long nblocks, xp_csr** out);
/* Host-side SELL build for inspection and CPU tests (no GPU): returns the
 * arrays of the format described in the library's xp_types.hpp.  With
 * block_rows = 8 and `slots` (units x 8 rows, -1 padding) the blocks are
 * kept together as in a block matrix; slots = NULL: 8 consecutive rows. */
int xp_sell_build_host(int n, const int* row_ptr, const int* col_idx, const double* values, int sigma,
                       int keep_lo, int keep64, int block_rows, const int* slots, int units, int* n_pad,
                       int* nslices, long long* stored, int* idx16);
/* The arrays of the last xp_sell_build_host of this thread (any may be NULL). */
int xp_sell_copy_host(int* perm, int* slice_ptr, int* slice_base, float* val, float* val_lo, short* val_lo16,
                      double* val64, unsigned short* col16, int* col32);

#ifdef __cplusplus
}
#endif

#endif /* XPLUS_CAPI_H */

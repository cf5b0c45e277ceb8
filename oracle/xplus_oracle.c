/*
 * xplus_oracle.c -- CPU restatement of the reference Jacobi-PCG path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product library (project-xplus_amd/) never links, loads or calls it, and
 * has no CPU fallback of any kind.
 *
 * Everything is plain C99 in FP64, single-threaded, in the same loop order
 * as the reference so that results can be compared bit for bit against the
 * reference build in oracle/_ref (see oracle/Makefile).  Each function cites
 * the reference file:line it restates.
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   orc_dataset_load        -- hand-written tiny files, one per error message;
 *                              equality with the reference loader (oracle/_ref).
 *   orc_spmv_csr            -- dense matvec on integer-valued inputs (exact);
 *                              bitwise equality with reference CsrDataset::spmv.
 *   orc_extract_jacobi_diag -- closed form + missing/zero-diagonal errors.
 *   orc_run_jacobi_pcg      -- bitwise equality with reference run_jacobi_pcg;
 *                              diagonal A converges in one iteration to b/diag;
 *                              convergence in <= n iterations for small n;
 *                              x against dense Cholesky (numpy); the n512
 *                              known answers from the reference's logs.
 *   orc_compute_residual_norm -- closed form on small inputs; reference equality.
 *   orc_k_* (six kernels)   -- bitwise equality with the reference
 *                              archived/kernels/cg_kernels.cpp for n <= 1024,
 *                              elementwise closed forms beyond.
 *   orc_multi_kernel_solver -- bitwise equality with the reference
 *                              run_local_multi_kernel_solver (n <= 1024).
 *   orc_cuper_style_spmv    -- bitwise equality with reference cuper_style_spmv.
 *   orc_jacobi_prepare_f32, orc_jacobi_run_f32 -- the reference's Jacobi host
 *                              (DLC/Cuper-jacobi-iteration/host/main.cpp)
 *                              needs TAPA/XRT and cannot be built here; pinned
 *                              by the closed forms in tests/test_jacobi.py:
 *                              one update of a diagonal system, the error
 *                              recursion e_k = (-D^-1 R)^k e_0 in FP64, the
 *                              fixed point A^-1 b, brute-force diff_max,
 *                              duplicate-diagonal sums, the A * ones rows.
 */
#include <math.h>
#include <float.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ORC_BREAKDOWN_EPS 1.0e-30 /* include/cg_common.hpp:13 kBreakdownEps */

typedef struct {
    int iteration;
    double alpha;
    double beta;
    double rz;
    double rr;
    double residual;
} orc_trace_t; /* src/CgSolverGolden.hpp:15-22 GoldenIterationTrace */

typedef struct {
    int n;
    int nnz;
    int* row_ptr;
    int* col_idx;
    double* values;
    double* b;
    double* x0;
} orc_dataset_t;

static void set_err(char* err, int errlen, const char* msg) {
    if (err != NULL && errlen > 0) {
        snprintf(err, (size_t)errlen, "%s", msg);
    }
}

/* Whitespace-separated numbers, stopping silently at the first token that
 * does not parse (the behaviour of `while (input >> value)`), error on an
 * unopenable or empty file.  src/CsrDataset.hpp:119-137 */
static int read_doubles(const char* path, double** out, long* count, char* err, int errlen) {
    FILE* f = fopen(path, "r");
    char msg[4096];
    if (f == NULL) {
        snprintf(msg, sizeof msg, "failed to open %s", path);
        set_err(err, errlen, msg);
        return -1;
    }
    long cap = 1024, len = 0;
    double* buf = (double*)malloc((size_t)cap * sizeof(double));
    double v;
    while (fscanf(f, "%lf", &v) == 1) {
        if (len == cap) {
            cap *= 2;
            buf = (double*)realloc(buf, (size_t)cap * sizeof(double));
        }
        buf[len++] = v;
    }
    fclose(f);
    if (len == 0) {
        free(buf);
        snprintf(msg, sizeof msg, "empty array file: %s", path);
        set_err(err, errlen, msg);
        return -1;
    }
    *out = buf;
    *count = len;
    return 0;
}

static int read_ints(const char* path, int** out, long* count, char* err, int errlen) {
    FILE* f = fopen(path, "r");
    char msg[4096];
    if (f == NULL) {
        snprintf(msg, sizeof msg, "failed to open %s", path);
        set_err(err, errlen, msg);
        return -1;
    }
    long cap = 1024, len = 0;
    int* buf = (int*)malloc((size_t)cap * sizeof(int));
    int v;
    while (fscanf(f, "%d", &v) == 1) {
        if (len == cap) {
            cap *= 2;
            buf = (int*)realloc(buf, (size_t)cap * sizeof(int));
        }
        buf[len++] = v;
    }
    fclose(f);
    if (len == 0) {
        free(buf);
        snprintf(msg, sizeof msg, "empty array file: %s", path);
        set_err(err, errlen, msg);
        return -1;
    }
    *out = buf;
    *count = len;
    return 0;
}

void orc_dataset_free(orc_dataset_t* d) {
    if (d == NULL) return;
    free(d->row_ptr);
    free(d->col_idx);
    free(d->values);
    free(d->b);
    free(d->x0);
    free(d);
}

/* CsrDataset::load, src/CsrDataset.hpp:19-67: same file order, same checks in
 * the same order, same messages. */
int orc_dataset_load(const char* dir, orc_dataset_t** out, char* err, int errlen) {
    char path[4096];
    long n_rp = 0, n_ci = 0, n_v = 0, n_b = 0, n_x = 0;
    orc_dataset_t* d = (orc_dataset_t*)calloc(1, sizeof(orc_dataset_t));
    *out = NULL;

    snprintf(path, sizeof path, "%s/row_ptr.txt", dir);
    if (read_ints(path, &d->row_ptr, &n_rp, err, errlen) != 0) goto fail;
    snprintf(path, sizeof path, "%s/col_idx.txt", dir);
    if (read_ints(path, &d->col_idx, &n_ci, err, errlen) != 0) goto fail;
    snprintf(path, sizeof path, "%s/values.txt", dir);
    if (read_doubles(path, &d->values, &n_v, err, errlen) != 0) goto fail;

    if (n_rp < 2) { set_err(err, errlen, "row_ptr.txt must contain at least 2 entries"); goto fail; }
    d->n = (int)(n_rp - 1);
    d->nnz = (int)n_ci;
    if (n_v != n_ci) { set_err(err, errlen, "col_idx.txt and values.txt size mismatch"); goto fail; }
    if (d->row_ptr[0] != 0) { set_err(err, errlen, "row_ptr[0] must be 0"); goto fail; }
    if (d->row_ptr[d->n] != d->nnz) { set_err(err, errlen, "row_ptr[n] must equal nnz"); goto fail; }
    for (int row = 0; row < d->n; ++row) {
        const int rb = d->row_ptr[row], re = d->row_ptr[row + 1];
        if (rb > re) { set_err(err, errlen, "row_ptr must be nondecreasing"); goto fail; }
        for (int o = rb; o < re; ++o) {
            const int c = d->col_idx[o];
            if (c < 0 || c >= d->n) { set_err(err, errlen, "col_idx out of range"); goto fail; }
        }
    }
    snprintf(path, sizeof path, "%s/b.txt", dir);
    if (read_doubles(path, &d->b, &n_b, err, errlen) != 0) goto fail;
    snprintf(path, sizeof path, "%s/x0.txt", dir);
    if (read_doubles(path, &d->x0, &n_x, err, errlen) != 0) goto fail;
    if ((int)n_b != d->n) { set_err(err, errlen, "b.txt length must equal n"); goto fail; }
    if ((int)n_x != d->n) { set_err(err, errlen, "x0.txt length must equal n"); goto fail; }
    *out = d;
    return 0;
fail:
    orc_dataset_free(d);
    return -1;
}

/* CsrDataset::spmv, src/CsrDataset.hpp:78-90: y[row] = sum in CSR order. */
void orc_spmv_csr(int n, const int* row_ptr, const int* col_idx, const double* values,
                  const double* x, double* y) {
    for (int row = 0; row < n; ++row) {
        double acc = 0.0;
        for (int o = row_ptr[row]; o < row_ptr[row + 1]; ++o) {
            acc += values[o] * x[col_idx[o]];
        }
        y[row] = acc;
    }
}

/* CsrDataset::extract_jacobi_diag, src/CsrDataset.hpp:92-116: first entry
 * with col == row; error if missing or |d| <= DBL_MIN. */
int orc_extract_jacobi_diag(int n, const int* row_ptr, const int* col_idx, const double* values,
                            double* diag, char* err, int errlen) {
    char msg[128];
    for (int row = 0; row < n; ++row) {
        int found = 0;
        diag[row] = 0.0;
        for (int o = row_ptr[row]; o < row_ptr[row + 1]; ++o) {
            if (col_idx[o] == row) {
                diag[row] = values[o];
                found = 1;
                break;
            }
        }
        if (!found) {
            snprintf(msg, sizeof msg, "missing diagonal entry at row %d", row);
            set_err(err, errlen, msg);
            return -1;
        }
        if (fabs(diag[row]) <= DBL_MIN) {
            snprintf(msg, sizeof msg, "zero diagonal entry at row %d", row);
            set_err(err, errlen, msg);
            return -1;
        }
    }
    return 0;
}

static double seq_dot(int n, const double* a, const double* b) {
    double acc = 0.0; /* src/CsrDataset.hpp:148-154 dot */
    for (int i = 0; i < n; ++i) acc += a[i] * b[i];
    return acc;
}

/* run_jacobi_pcg, src/CgSolverGolden.hpp:42-124.
 * Loop order and stopping rule as the reference: rr = r^T r <= tau on the
 * recursively updated residual, z = r / diag (a divide), beta computed and
 * traced on every iteration, break before the p update.  Returns 0, or -1
 * with the reference's message on tau <= 0, a bad diagonal or pAp breakdown.
 * trace may be NULL; at most trace_cap entries are written and *trace_len
 * receives the number of iterations traced. */
int orc_run_jacobi_pcg(int n, const int* row_ptr, const int* col_idx, const double* values,
                       const double* b, const double* x0, double tau, int max_iters_cfg,
                       double* x_out, orc_trace_t* trace, int trace_cap, int* trace_len,
                       int* iterations, double* final_rr, int* converged,
                       char* err, int errlen) {
    if (tau <= 0.0) {
        set_err(err, errlen, "tau must be positive");
        return -1;
    }
    const int max_iters = max_iters_cfg > 0 ? max_iters_cfg : (4 * n > 1000 ? 4 * n : 1000);
    double* diag = (double*)malloc((size_t)n * sizeof(double) + 8);
    if (orc_extract_jacobi_diag(n, row_ptr, col_idx, values, diag, err, errlen) != 0) {
        free(diag);
        return -1;
    }
    double* r = (double*)malloc((size_t)n * sizeof(double) + 8);
    double* z = (double*)malloc((size_t)n * sizeof(double) + 8);
    double* p = (double*)malloc((size_t)n * sizeof(double) + 8);
    double* ap = (double*)malloc((size_t)n * sizeof(double) + 8);
    int rc = 0, traced = 0;
    memcpy(x_out, x0, (size_t)n * sizeof(double));

    orc_spmv_csr(n, row_ptr, col_idx, values, x_out, ap);
    for (int i = 0; i < n; ++i) r[i] = b[i] - ap[i];
    for (int i = 0; i < n; ++i) z[i] = r[i] / diag[i];
    memcpy(p, z, (size_t)n * sizeof(double));
    double rz = seq_dot(n, r, z);
    double rr = seq_dot(n, r, r);
    *iterations = 0;
    *final_rr = rr;

    for (int it = 0; it < max_iters && rr > tau; ++it) {
        orc_spmv_csr(n, row_ptr, col_idx, values, p, ap);
        const double pap = seq_dot(n, p, ap);
        if (fabs(pap) <= DBL_MIN) {
            set_err(err, errlen, "breakdown: p^T A p is too small");
            rc = -1;
            break;
        }
        const double alpha = rz / pap;
        for (int i = 0; i < n; ++i) {
            x_out[i] += alpha * p[i];
            r[i] -= alpha * ap[i];
        }
        for (int i = 0; i < n; ++i) z[i] = r[i] / diag[i];
        const double rz_new = seq_dot(n, r, z);
        const double rr_new = seq_dot(n, r, r);
        const double beta = rz_new / rz;
        if (trace != NULL && traced < trace_cap) {
            trace[traced].iteration = it;
            trace[traced].alpha = alpha;
            trace[traced].beta = beta;
            trace[traced].rz = rz_new;
            trace[traced].rr = rr_new;
            trace[traced].residual = sqrt(rr_new > 0.0 ? rr_new : 0.0);
        }
        ++traced;
        rz = rz_new;
        rr = rr_new;
        *iterations = it + 1;
        *final_rr = rr;
        if (rr <= tau) break;
        for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
    if (trace_len != NULL) *trace_len = traced;
    *converged = (*final_rr <= tau) ? 1 : 0;
    free(diag); free(r); free(z); free(p); free(ap);
    return rc;
}

/* compute_residual_norm, src/CgSolverGolden.hpp:126-135: ||A x - b||_2. */
double orc_compute_residual_norm(int n, const int* row_ptr, const int* col_idx,
                                 const double* values, const double* b, const double* x) {
    double* ax = (double*)malloc((size_t)n * sizeof(double) + 8);
    orc_spmv_csr(n, row_ptr, col_idx, values, x, ax);
    for (int i = 0; i < n; ++i) ax[i] -= b[i];
    const double nrm = sqrt(seq_dot(n, ax, ax));
    free(ax);
    return nrm;
}

/* ---- the six fused operations, archived/kernels/cg_kernels.cpp:43-158,
 *      without the kMaxN cap (negative n is a silent no-op). ---- */

void orc_k_spmv_csr(const int* row_ptr, const int* col_idx, const double* values,
                    const double* x, double* y, int n) {
    if (n < 0) return;
    orc_spmv_csr(n, row_ptr, col_idx, values, x, y); /* :43-56 */
}

void orc_k_init_pcg(const double* b, const double* ax, const double* m_inv, double* r,
                    double* z, double* p, double* metrics, int n) {
    if (n < 0) return; /* :73-103 */
    double rz = 0.0, rr = 0.0;
    for (int i = 0; i < n; ++i) {
        const double rv = b[i] - ax[i];
        const double zv = m_inv[i] * rv;
        r[i] = rv;
        z[i] = zv;
        p[i] = zv;
        rz += rv * zv;
        rr += rv * rv;
    }
    metrics[0] = rz;
    metrics[1] = rr;
}

void orc_k_dot(const double* a, const double* b, double* out, int n) {
    if (n < 0) return; /* :105-116 */
    double acc = 0.0;
    for (int i = 0; i < n; ++i) acc += a[i] * b[i];
    out[0] = acc;
}

void orc_k_update_xrz(double* x, const double* p, double* r, const double* ap,
                      const double* m_inv, double* z, double* metrics, double alpha, int n) {
    if (n < 0) return; /* :118-147 */
    double rz = 0.0, rr = 0.0;
    for (int i = 0; i < n; ++i) {
        x[i] += alpha * p[i];
        r[i] -= alpha * ap[i];
        z[i] = m_inv[i] * r[i];
        rz += r[i] * z[i];
        rr += r[i] * r[i];
    }
    metrics[0] = rz;
    metrics[1] = rr;
}

void orc_k_update_p(const double* z, double* p, double beta, int n) {
    if (n < 0) return; /* :149-158 */
    for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
}

/* ---- run_local_multi_kernel_solver, archived/host/multi_kernel_solver.hpp:100-234 ---- */

enum { ORC_CONVERGED = 0, ORC_MAX_ITER = 1, ORC_BREAKDOWN = 2 };

static int finite(double v) { return isfinite(v) ? 1 : 0; }

/* Status semantics as the reference: breakdown (a status, not an error) on a
 * non-finite or tiny pAp / rz, non-finite alpha, rz_new, rr_new or beta;
 * update_p is called on every completed iteration including the converging
 * one; the true residual is taken at the end.  Returns -1 only where the
 * reference throws (tau <= 0, bad diagonal).  No kMaxN limit.
 * out_scalars = {final_rr, residual_l2, residual_rel}. */
int orc_multi_kernel_solver(int n, const int* row_ptr, const int* col_idx, const double* values,
                            const double* b, const double* x0, double tau, int max_iters_cfg,
                            double* x_out, orc_trace_t* trace, int trace_cap, int* iterations,
                            int* status, int* effective_max_iters, double* out_scalars,
                            char* err, int errlen) {
    if (tau <= 0.0) {
        set_err(err, errlen, "tau must be positive");
        return -1;
    }
    double* diag = (double*)malloc((size_t)n * sizeof(double) + 8);
    if (orc_extract_jacobi_diag(n, row_ptr, col_idx, values, diag, err, errlen) != 0) {
        free(diag);
        return -1;
    }
    double* m_inv = diag; /* build_jacobi_inverse :63-75 */
    for (int i = 0; i < n; ++i) {
        if (fabs(diag[i]) <= ORC_BREAKDOWN_EPS) {
            set_err(err, errlen, "zero diagonal entry while building Jacobi inverse");
            free(diag);
            return -1;
        }
        m_inv[i] = 1.0 / diag[i];
    }
    const int max_iters = max_iters_cfg > 0 ? max_iters_cfg : (4 * n > 1000 ? 4 * n : 1000);
    *effective_max_iters = max_iters;
    double* r = (double*)calloc((size_t)n + 1, sizeof(double));
    double* z = (double*)calloc((size_t)n + 1, sizeof(double));
    double* p = (double*)calloc((size_t)n + 1, sizeof(double));
    double* y = (double*)calloc((size_t)n + 1, sizeof(double));
    double metrics[2] = {0.0, 0.0}, dot_out[1] = {0.0};
    int st = ORC_MAX_ITER;
    memcpy(x_out, x0, (size_t)n * sizeof(double));
    *iterations = 0;

    orc_k_spmv_csr(row_ptr, col_idx, values, x_out, y, n);
    orc_k_init_pcg(b, y, m_inv, r, z, p, metrics, n);
    double rz = metrics[0], rr = metrics[1];
    double final_rr = rr;
    for (int it = 0; it < max_iters && rr > tau; ++it) {
        orc_k_spmv_csr(row_ptr, col_idx, values, p, y, n);
        orc_k_dot(p, y, dot_out, n);
        const double pap = dot_out[0];
        if (!finite(pap) || fabs(pap) <= ORC_BREAKDOWN_EPS || !finite(rz) ||
            fabs(rz) <= ORC_BREAKDOWN_EPS) {
            st = ORC_BREAKDOWN;
            break;
        }
        const double alpha = rz / pap;
        if (!finite(alpha)) { st = ORC_BREAKDOWN; break; }
        orc_k_update_xrz(x_out, p, r, y, m_inv, z, metrics, alpha, n);
        const double rz_new = metrics[0], rr_new = metrics[1];
        if (!finite(rz_new) || !finite(rr_new)) { st = ORC_BREAKDOWN; break; }
        const double beta = rz_new / rz;
        if (!finite(beta)) { st = ORC_BREAKDOWN; break; }
        orc_k_update_p(z, p, beta, n);
        rz = rz_new;
        rr = rr_new;
        *iterations = it + 1;
        final_rr = rr;
        if (trace != NULL && it < trace_cap) {
            trace[it].iteration = it;
            trace[it].alpha = alpha;
            trace[it].beta = beta;
            trace[it].rz = rz;
            trace[it].rr = rr;
            trace[it].residual = sqrt(rr > 0.0 ? rr : 0.0);
        }
    }
    /* true residual with the dataset's FP64 CSR SpMV, :217-226 */
    orc_spmv_csr(n, row_ptr, col_idx, values, x_out, y);
    for (int i = 0; i < n; ++i) y[i] -= b[i];
    const double res_l2 = sqrt(seq_dot(n, y, y));
    const double bn = sqrt(seq_dot(n, b, b));
    out_scalars[0] = final_rr;
    out_scalars[1] = res_l2;
    out_scalars[2] = res_l2 / (bn > ORC_BREAKDOWN_EPS ? bn : ORC_BREAKDOWN_EPS);
    if (st != ORC_BREAKDOWN) st = (final_rr <= tau) ? ORC_CONVERGED : ORC_MAX_ITER;
    *status = st;
    free(diag); free(r); free(z); free(p); free(y);
    return 0;
}

/* ---- cuper_style_spmv, host/cuper_pcg_solver.hpp:90-159 ----
 * Elements grouped into 8192-column batches (slice 64 x batch 128), in CSR
 * order inside each batch; x window cast to float; product float*float,
 * accumulated into double y in batch-major order. */
void orc_cuper_style_spmv(int n, const int* row_ptr, const int* col_idx, const double* values,
                          const double* x, double* y) {
    const int slice_width = 8192;
    const int num_batches = ((n + 63) / 64 + 127) / 128;
    const int nnz = row_ptr[n];
    int* start = (int*)calloc((size_t)num_batches + 1, sizeof(int));
    int* order = (int*)malloc((size_t)nnz * sizeof(int) + 4);
    int* rows = (int*)malloc((size_t)nnz * sizeof(int) + 4);
    for (int o = 0; o < nnz; ++o) start[col_idx[o] / slice_width + 1]++;
    for (int bt = 0; bt < num_batches; ++bt) start[bt + 1] += start[bt];
    for (int row = 0; row < n; ++row) {
        for (int o = row_ptr[row]; o < row_ptr[row + 1]; ++o) {
            const int slot = start[col_idx[o] / slice_width]++;
            order[slot] = o;
            rows[slot] = row;
        }
    }
    for (int i = 0; i < n; ++i) y[i] = 0.0;
    for (int e = 0; e < nnz; ++e) {
        const int o = order[e];
        const float xv = (float)x[col_idx[o]];
        const float prod = (float)values[o] * xv;
        y[rows[e]] += (double)prod;
    }
    free(start); free(order); free(rows);
}

/* ---- Plain Jacobi iteration, FP32 (VALUE_TYPE float), ----
 * DLC/Cuper-jacobi-iteration/host/main.cpp:957-1009 PrepareJacobiVectorsAndR:
 * values rounded to float; without a b, b = A * ones (SpMV_CPU_CSR,
 * include/Cuper_common.h:299-320: float sum from 0 in CSR order); D = float
 * sum of the row's diagonal entries in CSR order; R = the other entries in
 * CSR order; Diag_inv = 1.0f / D.  Returns -1, or the first row with a
 * missing or zero diagonal. */
int orc_jacobi_prepare_f32(int n, const int* row_ptr, const int* col_idx, const double* values, const double* b_in,
                           float* b, int* r_rp, int* r_ci, float* r_val, float* dinv) {
    for (int row = 0; row < n; ++row) {
        if (b_in != NULL) {
            b[row] = (float)b_in[row];
        } else {
            float sum = 0.0f;
            for (int o = row_ptr[row]; o < row_ptr[row + 1]; ++o) sum += 1.0f * (float)values[o];
            b[row] = sum;
        }
    }
    int k = 0;
    for (int row = 0; row < n; ++row) {
        int found = 0;
        float d = 0.0f;
        r_rp[row] = k;
        for (int o = row_ptr[row]; o < row_ptr[row + 1]; ++o) {
            if (col_idx[o] == row) {
                d += (float)values[o];
                found = 1;
            } else {
                r_ci[k] = col_idx[o];
                r_val[k] = (float)values[o];
                ++k;
            }
        }
        if (!found || d == 0.0f) return row;
        dinv[row] = 1.0f / d;
    }
    r_rp[n] = k;
    return -1;
}

/* RunJacobiCpu (host/main.cpp:1011-1053): every update forms x_next from the
 * whole old x (R x in float, CSR order, from 0), next = (b - rx) * Diag_inv,
 * diff_max = max |next - x|, then swaps.  Without use_tau exactly max_iters
 * updates (the reference's fixed count), status breakdown when the last
 * difference is NaN or infinite, else max_iter.  With use_tau the loop also
 * ends once diff_max <= tau (docs/jacobi_iteration.md:99-124), and a NaN or
 * infinite difference ends it as breakdown (:148-154).
 * x holds x0 on entry and the result on return; returns the update count. */
int orc_jacobi_run_f32(int n, const int* r_rp, const int* r_ci, const float* r_val, const float* b, const float* dinv,
                       int max_iters, int use_tau, double tau, float* x, float* final_diff, int* status) {
    float* nx = (float*)malloc((size_t)(n > 0 ? n : 1) * sizeof(float));
    int it = 0;
    *status = ORC_MAX_ITER;
    *final_diff = 0.0f;
    while (it < max_iters) {
        float dmax = 0.0f;
        int nan_seen = 0;
        for (int row = 0; row < n; ++row) {
            float rx = 0.0f;
            for (int o = r_rp[row]; o < r_rp[row + 1]; ++o) rx += x[r_ci[o]] * r_val[o];
            const float next = (b[row] - rx) * dinv[row];
            const float diff = fabsf(next - x[row]);
            nx[row] = next;
            if (diff > dmax) dmax = diff;
            if (isnan(diff)) nan_seen = 1;
        }
        memcpy(x, nx, (size_t)n * sizeof(float));
        ++it;
        *final_diff = nan_seen ? NAN : dmax;   /* a NaN difference is reported as NaN */
        if (!use_tau) continue;                /* fixed count: every update runs */
        if (nan_seen || isinf(dmax)) { *status = ORC_BREAKDOWN; break; }
        if ((double)dmax <= tau) { *status = ORC_CONVERGED; break; }
    }
    if (!use_tau && it > 0 && !isfinite(*final_diff)) *status = ORC_BREAKDOWN;
    free(nx);
    return it;
}

void orc_free(void* ptr) { free(ptr); }

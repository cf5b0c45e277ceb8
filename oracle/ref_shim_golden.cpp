// ref_shim_golden.cpp -- C entry points over the reference's own header-only
// golden, loader and multi-kernel orchestrator, compiled by oracle/Makefile
// from the sources under the reference tree (nothing copied).  TEST
// INFRASTRUCTURE ONLY: used by tests/ to pin oracle/xplus_oracle.c and, when
// present, by bench.py as the "reference" CPU baseline.
//
// Wraps: CsrDataset::load/spmv/extract_jacobi_diag (src/CsrDataset.hpp:19-116),
// run_cpu_reference -> run_jacobi_pcg + compute_residual_norm
// (host/cpu_reference.hpp:15-30, src/CgSolverGolden.hpp:42-135),
// run_local_multi_kernel_solver (archived/host/multi_kernel_solver.hpp:100-234).
#include "cpu_reference.hpp"
#include "multi_kernel_solver.hpp"

#include <chrono>
#include <cstdio>
#include <cstring>
#include <exception>

namespace pc = project_xplus::cgsolver;

namespace {
void put_err(char* err, int errlen, const char* msg) {
    if (err != nullptr && errlen > 0) std::snprintf(err, static_cast<size_t>(errlen), "%s", msg);
}
}  // namespace

extern "C" {

int ref_dataset_info(const char* dir, int* n, int* nnz, char* err, int errlen) {
    try {
        const pc::CsrDataset d = pc::CsrDataset::load(dir);
        *n = d.n();
        *nnz = d.nnz();
        return 0;
    } catch (const std::exception& e) {
        put_err(err, errlen, e.what());
        return -1;
    }
}

int ref_dataset_read(const char* dir, int* row_ptr, int* col_idx, double* values, double* b,
                     double* x0, char* err, int errlen) {
    try {
        const pc::CsrDataset d = pc::CsrDataset::load(dir);
        std::memcpy(row_ptr, d.row_ptr().data(), d.row_ptr().size() * sizeof(int));
        std::memcpy(col_idx, d.col_idx().data(), d.col_idx().size() * sizeof(int));
        std::memcpy(values, d.values().data(), d.values().size() * sizeof(double));
        std::memcpy(b, d.b().data(), d.b().size() * sizeof(double));
        std::memcpy(x0, d.x0().data(), d.x0().size() * sizeof(double));
        return 0;
    } catch (const std::exception& e) {
        put_err(err, errlen, e.what());
        return -1;
    }
}

int ref_dataset_spmv(const char* dir, const double* x, double* y, char* err, int errlen) {
    try {
        const pc::CsrDataset d = pc::CsrDataset::load(dir);
        std::vector<double> xv(x, x + d.n()), yv;
        d.spmv(xv, yv);
        std::memcpy(y, yv.data(), yv.size() * sizeof(double));
        return 0;
    } catch (const std::exception& e) {
        put_err(err, errlen, e.what());
        return -1;
    }
}

int ref_extract_jacobi_diag(const char* dir, double* diag, char* err, int errlen) {
    try {
        const pc::CsrDataset d = pc::CsrDataset::load(dir);
        const std::vector<double> dg = d.extract_jacobi_diag();
        std::memcpy(diag, dg.data(), dg.size() * sizeof(double));
        return 0;
    } catch (const std::exception& e) {
        put_err(err, errlen, e.what());
        return -1;
    }
}

// Runs run_cpu_reference.  trace rows are {iteration, alpha, beta, rz, rr,
// residual} packed as 6 doubles.  scalars = {final_rr, residual_l2, solve_ms}.
int ref_golden(const char* dir, double tau, int max_iters, double* x_out, double* trace,
               int trace_cap, int* iterations, int* converged, double* scalars, char* err,
               int errlen) {
    try {
        const pc::CsrDataset d = pc::CsrDataset::load(dir);
        pc::SolverConfig cfg;
        cfg.tau = tau;
        cfg.max_iters = max_iters;
        const auto t0 = std::chrono::steady_clock::now();
        const pc::CpuReferenceResult res = pc::run_cpu_reference(d, cfg);
        const auto t1 = std::chrono::steady_clock::now();
        std::memcpy(x_out, res.golden.solution.data(), res.golden.solution.size() * sizeof(double));
        const int m = std::min<int>(trace_cap, static_cast<int>(res.golden.iteration_trace.size()));
        for (int i = 0; i < m; ++i) {
            const auto& t = res.golden.iteration_trace[static_cast<size_t>(i)];
            trace[6 * i + 0] = t.iteration;
            trace[6 * i + 1] = t.alpha;
            trace[6 * i + 2] = t.beta;
            trace[6 * i + 3] = t.rz;
            trace[6 * i + 4] = t.rr;
            trace[6 * i + 5] = t.residual;
        }
        *iterations = res.golden.iterations;
        *converged = res.golden.converged ? 1 : 0;
        scalars[0] = res.golden.final_rr;
        scalars[1] = res.summary.residual_l2;
        scalars[2] = std::chrono::duration<double, std::milli>(t1 - t0).count();
        return 0;
    } catch (const std::exception& e) {
        put_err(err, errlen, e.what());
        return -1;
    }
}

// Runs run_local_multi_kernel_solver with the reference's own CPU kernels
// (archived/kernels/cg_kernels.cpp, linked into this library).
// scalars = {final_rr, residual_l2, residual_rel}; status 0/1/2.
int ref_multi_kernel(const char* dir, double tau, int max_iters, double* x_out, int* iterations,
                     int* status, int* effective_max_iters, double* scalars, char* err,
                     int errlen) {
    try {
        const pc::CsrDataset d = pc::CsrDataset::load(dir);
        pc::SolverConfig cfg;
        cfg.tau = tau;
        cfg.max_iters = max_iters;
        const pc::MultiKernelResult res = pc::run_local_multi_kernel_solver(d, cfg, nullptr);
        std::memcpy(x_out, res.solution.data(), res.solution.size() * sizeof(double));
        *iterations = res.iterations;
        *status = static_cast<int>(res.status);
        *effective_max_iters = res.effective_max_iters;
        scalars[0] = res.final_rr;
        scalars[1] = res.residual_l2;
        scalars[2] = res.residual_rel;
        return 0;
    } catch (const std::exception& e) {
        put_err(err, errlen, e.what());
        return -1;
    }
}

}  // extern "C"

// multi_kernel_solver.hpp -- host orchestration over the drop-in C-ABI.
//
// Restates run_local_multi_kernel_solver (archived/host/multi_kernel_solver.hpp:
// 100-234): the same loop (spmv -> dot -> host alpha -> update_xrz -> host
// beta -> update_p), the same breakdown tests (non-finite or |.| <= 1e-30
// pAp and rz, non-finite alpha, rz_new, rr_new, beta), update_p on every
// completed iteration, the same [init]/[iter NNN] log lines and the same
// final true residual.  The five operations are the library's GPU
// implementations of the reference names; the true residual uses the GPU
// FP64 SpMV (CsrDataset::spmv).  There is no n <= 1024 limit.
#pragma once

#include <algorithm>
#include <cmath>
#include <iomanip>
#include <ostream>
#include <stdexcept>
#include <vector>

#include "../../include/xplus_capi.h"
#include "csr_dataset.hpp"

namespace xp {

enum class SolverStatus { kConverged, kMaxIter, kBreakdown };

inline const char* to_string(SolverStatus s) {
    switch (s) {
        case SolverStatus::kConverged: return "converged";
        case SolverStatus::kMaxIter: return "max_iter";
        case SolverStatus::kBreakdown: return "breakdown";
    }
    return "unknown";
}

struct SolverConfig {
    double tau = 1.0e-10;  // include/cg_common.hpp:53
    int max_iters = 0;     // 0 -> max(4n, 1000)
};

struct MultiKernelResult {
    std::vector<double> solution;
    std::vector<double> jacobi_inverse;
    int iterations = 0;
    int effective_max_iters = 0;
    double final_rr = 0.0;
    double residual_l2 = 0.0;
    double residual_rel = 0.0;
    SolverStatus status = SolverStatus::kMaxIter;
    bool converged = false;
};

using SpmvKernelFn = void (*)(const int*, const int*, const double*, const double*, double*, int);

constexpr double kBreakdownEps = 1.0e-30;

inline double l2(const std::vector<double>& v) {
    double acc = 0.0;
    for (double e : v) acc += e * e;
    return std::sqrt(acc);
}

inline void log_iteration(std::ostream& out, int it, double alpha, double beta, double rz, double rr) {
    const auto flags = out.flags();
    const auto prec = out.precision();
    const char fill = out.fill();
    out << "[iter " << std::setw(3) << std::setfill('0') << it << "] " << std::scientific
        << std::setprecision(12) << "alpha=" << alpha << " beta=" << beta << " rz=" << rz << " rr=" << rr
        << " residual=" << std::sqrt(std::max(rr, 0.0)) << "\n";
    out.flags(flags);
    out.precision(prec);
    out.fill(fill);
}

inline MultiKernelResult run_local_multi_kernel_solver(const CsrDataset& d, const SolverConfig& cfg,
                                                       std::ostream* log = nullptr,
                                                       SpmvKernelFn spmv = spmv_csr_kernel) {
    if (cfg.tau <= 0.0) throw std::runtime_error("tau must be positive");
    MultiKernelResult res;
    const int n = d.n();
    res.solution = d.x0();
    {
        const std::vector<double> diag = d.extract_jacobi_diag();
        res.jacobi_inverse.assign(diag.size(), 0.0);
        for (size_t i = 0; i < diag.size(); ++i) {
            if (std::fabs(diag[i]) <= kBreakdownEps)
                throw std::runtime_error("zero diagonal entry while building Jacobi inverse");
            res.jacobi_inverse[i] = 1.0 / diag[i];
        }
    }
    res.effective_max_iters = cfg.max_iters > 0 ? cfg.max_iters : std::max(4 * n, 1000);
    std::vector<double> r(static_cast<size_t>(n), 0.0), z(r), p(r), y(r);
    double metrics[2] = {0.0, 0.0};
    double dot_out[1] = {0.0};
    if (log != nullptr) {
        *log << "[init] n=" << n << " nnz=" << d.nnz() << " tau=" << std::scientific << std::setprecision(12)
             << cfg.tau << " max_iters=" << std::defaultfloat << res.effective_max_iters << " dtype=double\n";
    }
    spmv(d.row_ptr().data(), d.col_idx().data(), d.values().data(), res.solution.data(), y.data(), n);
    init_pcg_kernel(d.b().data(), y.data(), res.jacobi_inverse.data(), r.data(), z.data(), p.data(), metrics, n);
    double rz = metrics[0], rr = metrics[1];
    res.final_rr = rr;
    for (int it = 0; it < res.effective_max_iters && rr > cfg.tau; ++it) {
        spmv(d.row_ptr().data(), d.col_idx().data(), d.values().data(), p.data(), y.data(), n);
        dot_kernel(p.data(), y.data(), dot_out, n);
        const double pap = dot_out[0];
        if (!std::isfinite(pap) || std::fabs(pap) <= kBreakdownEps || !std::isfinite(rz) ||
            std::fabs(rz) <= kBreakdownEps) {
            res.status = SolverStatus::kBreakdown;
            break;
        }
        const double alpha = rz / pap;
        if (!std::isfinite(alpha)) { res.status = SolverStatus::kBreakdown; break; }
        update_xrz_kernel(res.solution.data(), p.data(), r.data(), y.data(), res.jacobi_inverse.data(), z.data(),
                          metrics, alpha, n);
        const double rz_new = metrics[0], rr_new = metrics[1];
        if (!std::isfinite(rz_new) || !std::isfinite(rr_new)) { res.status = SolverStatus::kBreakdown; break; }
        const double beta = rz_new / rz;
        if (!std::isfinite(beta)) { res.status = SolverStatus::kBreakdown; break; }
        update_p_kernel(z.data(), p.data(), beta, n);
        rz = rz_new;
        rr = rr_new;
        res.iterations = it + 1;
        res.final_rr = rr;
        if (log != nullptr) log_iteration(*log, it, alpha, beta, rz, rr);
    }
    std::vector<double> ax;
    d.spmv(res.solution, ax);
    for (int i = 0; i < n; ++i) ax[static_cast<size_t>(i)] -= d.b()[static_cast<size_t>(i)];
    res.residual_l2 = l2(ax);
    res.residual_rel = res.residual_l2 / std::max(l2(d.b()), kBreakdownEps);
    res.converged = res.final_rr <= cfg.tau;
    if (res.status != SolverStatus::kBreakdown)
        res.status = res.converged ? SolverStatus::kConverged : SolverStatus::kMaxIter;
    return res;
}

}  // namespace xp

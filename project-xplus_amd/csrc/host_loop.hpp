// host_loop.hpp -- a PCG loop driven from the host over the five drop-in
// operations of the C-ABI (spmv, init_pcg, dot, update_xrz, update_p).
//
// Contract kept from the reference's local multi-kernel host
// (archived/host/multi_kernel_solver.hpp:100-234), because its callers and
// scripts depend on it: the operation order of a step (A p, p^T A p,
// alpha on the host, x/r/z update with rz and rr, beta on the host, p
// update, also on the step that converges), a breakdown whenever p^T A p or
// rz is non-finite or at most 1e-30 in size or alpha, the new rz / rr or
// beta is non-finite, the maximum step count max(4n, 1000) when none is
// given, and the "[init] ..." / "[iter NNN] ..." lines.  Everything else is
// this library's: the loop is a small state machine (PcgDriver) whose state
// lives in one struct, the checks are one predicate, and the final true
// residual comes from the GPU FP64 SpMV.  There is no n <= 1024 limit.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <ostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/xplus_capi.h"
#include "csr_dataset.hpp"

namespace xp {

constexpr double kBreakdownEps = 1.0e-30;  // include/cg_common.hpp:13

// How a host-driven solve ended; the names are the ones the reference prints.
enum class Outcome { converged, max_iter, breakdown };

inline const char* outcome_name(Outcome o) {
    return o == Outcome::converged ? "converged" : (o == Outcome::breakdown ? "breakdown" : "max_iter");
}

struct LoopSettings {
    double tau = 1.0e-10;  // bound on r^T r (include/cg_common.hpp:53)
    int max_iters = 0;     // 0: max(4 n, 1000)
};

struct LoopResult {
    std::vector<double> x;          // the solution
    std::vector<double> inv_diag;   // 1 / diag, as applied
    int steps = 0;                  // completed steps
    int step_cap = 0;               // the maximum used
    double rr = 0.0;                // r^T r after the last completed step (after init if none)
    double residual_l2 = 0.0;       // ||b - A x||
    double residual_rel = 0.0;      // ||b - A x|| / max(||b||, 1e-30)
    Outcome outcome = Outcome::max_iter;
    bool converged = false;
};

using SpmvEntry = void (*)(const int*, const int*, const double*, const double*, double*, int);

inline double norm2(const std::vector<double>& v) {
    double acc = 0.0;
    for (double e : v) acc += e * e;
    return std::sqrt(acc);
}

class PcgDriver {
  public:
    PcgDriver(const CsrDataset& d, const LoopSettings& s, SpmvEntry spmv) : d_(d), s_(s), spmv_(spmv) {
        if (!(s.tau > 0.0)) throw std::runtime_error("tau must be positive");
    }

    LoopResult solve(std::ostream* log) {
        LoopResult out;
        const int n = d_.n();
        prepare(out);
        if (log != nullptr) header(*log, out.step_cap);
        std::vector<double> r(static_cast<size_t>(n)), z(r), p(r), q(r);
        double m2[2] = {0.0, 0.0};
        apply(out.x, q);  // q = A x0
        init_pcg_kernel(d_.b().data(), q.data(), out.inv_diag.data(), r.data(), z.data(), p.data(), m2, n);
        double rz = m2[0];
        out.rr = m2[1];
        while (out.steps < out.step_cap && out.rr > s_.tau) {
            apply(p, q);  // q = A p
            double pq = 0.0;
            dot_kernel(p.data(), q.data(), &pq, n);
            if (degenerate(pq) || degenerate(rz)) return finish(out, Outcome::breakdown);
            const double alpha = rz / pq;
            if (!std::isfinite(alpha)) return finish(out, Outcome::breakdown);
            update_xrz_kernel(out.x.data(), p.data(), r.data(), q.data(), out.inv_diag.data(), z.data(), m2, alpha,
                              n);
            if (!std::isfinite(m2[0]) || !std::isfinite(m2[1])) return finish(out, Outcome::breakdown);
            const double beta = m2[0] / rz;
            if (!std::isfinite(beta)) return finish(out, Outcome::breakdown);
            update_p_kernel(z.data(), p.data(), beta, n);
            rz = m2[0];
            out.rr = m2[1];
            if (log != nullptr) step_line(*log, out.steps, alpha, beta, rz, out.rr);
            ++out.steps;
        }
        return finish(out, out.rr <= s_.tau ? Outcome::converged : Outcome::max_iter);
    }

  private:
    static bool degenerate(double v) { return !std::isfinite(v) || std::fabs(v) <= kBreakdownEps; }

    void apply(const std::vector<double>& v, std::vector<double>& av) const {
        spmv_(d_.row_ptr().data(), d_.col_idx().data(), d_.values().data(), v.data(), av.data(), d_.n());
    }

    void prepare(LoopResult& out) const {
        out.x = d_.x0();
        const std::vector<double> diag = d_.extract_jacobi_diag();
        out.inv_diag.resize(diag.size());
        for (size_t i = 0; i < diag.size(); ++i) {
            if (std::fabs(diag[i]) <= kBreakdownEps)
                throw std::runtime_error("zero diagonal entry while building Jacobi inverse");
            out.inv_diag[i] = 1.0 / diag[i];
        }
        out.step_cap = s_.max_iters > 0 ? s_.max_iters : std::max(4 * d_.n(), 1000);
    }

    // The true residual of the final x (GPU FP64 SpMV) and the outcome.
    LoopResult& finish(LoopResult& out, Outcome o) const {
        std::vector<double> ax;
        d_.spmv(out.x, ax);
        for (size_t i = 0; i < ax.size(); ++i) ax[i] -= d_.b()[i];
        out.residual_l2 = norm2(ax);
        out.residual_rel = out.residual_l2 / std::max(norm2(d_.b()), kBreakdownEps);
        out.converged = out.rr <= s_.tau;
        out.outcome = o == Outcome::breakdown ? o : (out.converged ? Outcome::converged : Outcome::max_iter);
        return out;
    }

    void header(std::ostream& os, int cap) const {
        char buf[160];
        std::snprintf(buf, sizeof buf, "[init] n=%d nnz=%d tau=%.12e max_iters=%d dtype=double\n", d_.n(), d_.nnz(),
                      s_.tau, cap);
        os << buf;
    }

    static void step_line(std::ostream& os, int it, double alpha, double beta, double rz, double rr) {
        char buf[200];
        std::snprintf(buf, sizeof buf, "[iter %03d] alpha=%.12e beta=%.12e rz=%.12e rr=%.12e residual=%.12e\n", it,
                      alpha, beta, rz, rr, std::sqrt(std::max(rr, 0.0)));
        os << buf;
    }

    const CsrDataset& d_;
    LoopSettings s_;
    SpmvEntry spmv_;
};

inline LoopResult run_host_pcg(const CsrDataset& d, const LoopSettings& s, std::ostream* log = nullptr,
                               SpmvEntry spmv = spmv_csr_kernel) {
    return PcgDriver(d, s, spmv).solve(log);
}

}  // namespace xp

// ref_cuper_backend_on_amd.cpp -- the reference's pluggable-SpMV PCG loop
// (run_cuper_pcg_solver_with_backend, host/cuper_pcg_solver.hpp:169-309),
// compiled unchanged from the reference tree by oracle/Makefile, with
// libxplus_amd.so as its SpmvRunner.  This is how a reference host
// (host/cuper_tapa_pcg_main.cpp:390-418 binds CuperTapaSpmv) would plug the
// MI355X SpMV in; the pass rule is that host's own (max_abs_diff or
// max_rel_diff against run_cpu_reference within diff_tol).
//
//   cuper_pcg_on_amd <dataset dir> [--backend csr64|sell32] [--tau T]
//                    [--max-iters N] [--diff-tol D]
//
// csr64:  spmv_csr_kernel, the drop-in FP64 CSR entry point (the matrix is
//         converted and cached on the device on the first call);
// sell32: the device SELL matrix and the FP32 SpMV (xp_spmv_f32), the
//         vectors moved to and from solve order on the device -- the
//         reference's own FP32-SpMV / FP64-PCG split.
// Exit codes: 0 pass, 3 diff above diff_tol, 1 error.  Built only where the
// reference tree is present; the binary travels with the repository tree.
#include "cpu_reference.hpp"
#include "cuper_pcg_solver.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "xplus_capi.h"

namespace pc = project_xplus::cgsolver;

namespace {

void check(int rc, const char* what) {
    if (rc != 0) {
        char buf[512];
        xp_last_error(buf, sizeof buf);
        throw std::runtime_error(std::string(what) + ": " + buf);
    }
}

// SpmvRunner over the library: operator()(input, output, result) as the
// reference's CuperTapaSpmv provides it.
class AmdSpmv {
public:
    AmdSpmv(const pc::Dataset& d, bool sell32) : d_(d), sell32_(sell32) {
        const auto t0 = std::chrono::steady_clock::now();
        if (sell32_) {
            xp_matrix_opts o{};
            o.sigma = 256;
            check(xp_matrix_create(d.n(), d.row_ptr().data(), d.col_idx().data(), d.values().data(), &o, &m_),
                  "xp_matrix_create");
            xp_matrix_info info{};
            check(xp_matrix_get_info(m_, &info), "xp_matrix_get_info");
            const unsigned long long n = static_cast<unsigned long long>(d.n());
            const unsigned long long np = static_cast<unsigned long long>(info.n_pad);
            check(xp_device_alloc(n * sizeof(double), &nat_), "xp_device_alloc");
            check(xp_device_alloc(np * sizeof(float), &xs_), "xp_device_alloc");
            check(xp_device_alloc(np * sizeof(float), &ys_), "xp_device_alloc");
        }
        plan_ms_ = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    AmdSpmv(const AmdSpmv&) = delete;
    AmdSpmv& operator=(const AmdSpmv&) = delete;
    ~AmdSpmv() {
        if (nat_ != nullptr) xp_device_free(nat_);
        if (xs_ != nullptr) xp_device_free(xs_);
        if (ys_ != nullptr) xp_device_free(ys_);
        if (m_ != nullptr) xp_matrix_destroy(m_);
    }

    pc::CuperPcgBackendInfo backend_info() const {
        return pc::CuperPcgBackendInfo{sell32_ ? "mi355x-sell-fp32-spmv+fp64-pcg" : "mi355x-csr-fp64-spmv+fp64-pcg",
                                       0, 0};
    }
    double plan_ms() const { return plan_ms_; }

    void operator()(const std::vector<pc::data_t>& input, std::vector<pc::data_t>& output, pc::CuperPcgResult& result) {
        const auto t0 = std::chrono::steady_clock::now();
        output.assign(static_cast<std::size_t>(d_.n()), 0.0);
        if (sell32_) {
            const unsigned long long bytes = static_cast<unsigned long long>(d_.n()) * sizeof(double);
            check(xp_memcpy_htod(nat_, input.data(), bytes), "xp_memcpy_htod");
            check(xp_to_solve_order(m_, static_cast<const double*>(nat_), static_cast<float*>(xs_), nullptr),
                  "xp_to_solve_order");
            check(xp_spmv_f32(m_, static_cast<const float*>(xs_), static_cast<float*>(ys_), nullptr), "xp_spmv_f32");
            check(xp_from_solve_order(m_, static_cast<const float*>(ys_), static_cast<double*>(nat_), nullptr),
                  "xp_from_solve_order");
            check(xp_memcpy_dtoh(output.data(), nat_, bytes), "xp_memcpy_dtoh");
        } else {
            spmv_csr_kernel(d_.row_ptr().data(), d_.col_idx().data(), d_.values().data(), input.data(),
                            output.data(), d_.n());
            char buf[8];
            if (xp_last_error(buf, sizeof buf) > 0) check(-1, "spmv_csr_kernel");
        }
        result.timing.spmv_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        ++result.timing.spmv_calls;
    }

private:
    const pc::Dataset& d_;
    bool sell32_;
    double plan_ms_ = 0.0;
    xp_matrix* m_ = nullptr;
    void* nat_ = nullptr;
    void* xs_ = nullptr;
    void* ys_ = nullptr;
};

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::cerr << "usage: " << argv[0] << " <dataset dir> [--backend csr64|sell32] [--tau T] [--max-iters N]"
                  << " [--diff-tol D]\n";
        return 1;
    }
    std::string dir = argv[1], backend = "csr64";
    pc::SolverConfig config;
    double diff_tol = 1.0e-3;  // host/cuper_tapa_pcg_main.cpp:47
    for (int i = 2; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--backend" && i + 1 < argc) backend = argv[++i];
        else if (a == "--tau" && i + 1 < argc) config.tau = std::stod(argv[++i]);
        else if (a == "--max-iters" && i + 1 < argc) config.max_iters = std::stoi(argv[++i]);
        else if (a == "--diff-tol" && i + 1 < argc) diff_tol = std::stod(argv[++i]);
        else {
            std::cerr << "unknown option " << a << "\n";
            return 1;
        }
    }
    if (backend != "csr64" && backend != "sell32") {
        std::cerr << "--backend must be csr64 or sell32\n";
        return 1;
    }
    try {
        const pc::Dataset dataset = pc::Dataset::load(dir);
        AmdSpmv spmv(dataset, backend == "sell32");
        const pc::CuperPcgResult r =
            pc::run_cuper_pcg_solver_with_backend(dataset, config, spmv.backend_info(), spmv.plan_ms(), std::ref(spmv),
                                                  &std::cout);
        const pc::CpuReferenceResult golden = pc::run_cpu_reference(dataset, config);
        double max_abs_diff = 0.0, max_rel_diff = 0.0;
        for (std::size_t i = 0; i < r.solution.size(); ++i) {
            const double e = golden.golden.solution[i];
            const double d = std::fabs(r.solution[i] - e);
            max_abs_diff = std::max(max_abs_diff, d);
            max_rel_diff = std::max(max_rel_diff, d / std::max(std::fabs(e), 1.0e-12));
        }
        std::cout << std::scientific << std::setprecision(12);
        std::cout << "[done] iter=" << r.iterations << " residual_abs=" << r.residual_l2
                  << " residual_rel=" << r.residual_rel << " status=" << pc::to_string(r.status) << "\n";
        std::cout << "[check] golden_iter=" << golden.golden.iterations
                  << " cpu_residual_abs=" << golden.summary.residual_l2 << " amd_residual_abs=" << r.residual_l2
                  << " max_abs_diff=" << max_abs_diff << " max_rel_diff=" << max_rel_diff
                  << " diff_tol=" << diff_tol << "\n";
        std::cout << "[timing-ms] plan=" << r.timing.plan_ms << " spmv_total=" << r.timing.spmv_ms
                  << " spmv_calls=" << std::defaultfloat << r.timing.spmv_calls << "\n";
        if (max_abs_diff > diff_tol && max_rel_diff > diff_tol) return 3;
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
}

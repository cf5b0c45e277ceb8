// xplus_cuper_main.cpp -- command-line driver with the reference's Cuper
// host command line (host/cuper_tapa_pcg_main.cpp:42-125):
//   [dataset_dir] [--bitstream path] [--tau v] [--max-iters v]
//   [--diff-tol v] [--spmv-only] [--pcg-spmv-service] [--spmv-repeats v]
// --bitstream / BITFILE and --pcg-spmv-service select FPGA images there and
// are accepted and reported here, with no effect.
//
// --spmv-only reproduces the reference's SpMV measurement
// (cuper_tapa_pcg_main.cpp:331-386): the matrix goes through the Cuper
// packed format (xp_cuper_pack, the reference's packer) and back
// (xp_cuper_unpack) into a device SELL matrix -- the plan, timed as
// plan_ms -- then y = A b in FP32 (the packed FP32 values, b rounded to
// FP32) spmv_repeats times, each launch timed with HIP events; the check is
// max |y - A b| against the FP64 product, absolute and relative, against
// diff_tol; GFLOP/s = 2 nnz calls / kernel seconds.  Same output lines, same
// exit codes (3: both differences above diff_tol).
//
// The default mode runs the graph-captured FP32 Jacobi-PCG on the device
// and checks its solution against the reference arithmetic of the FP64
// multi-kernel drop-in solver (the same role as the reference's CPU golden
// there, :392-440), with the same [done] / [check] / [timing-ms] lines.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <filesystem>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "host_loop.hpp"

namespace {

struct Cli {
    std::filesystem::path dir = "data/suitesparse/Schmid/csr/thermal2_n1024";
    std::string bitstream;
    double tau = 1.0e-10;
    int max_iters = 0;
    double diff_tol = 1.0e-3;
    bool spmv_only = false;
    bool pcg_spmv_service = false;
    int spmv_repeats = 1;
};

void usage(const char* argv0) {
    std::cerr << "Usage: " << argv0
              << " [dataset_dir] [--bitstream path] [--tau value] [--max-iters value]"
              << " [--diff-tol value] [--spmv-only] [--pcg-spmv-service]"
              << " [--spmv-repeats value]\n";
}

Cli parse(int argc, char** argv) {
    Cli c;
    if (const char* e = std::getenv("BITFILE")) c.bitstream = e;
    int i = 1;
    if (i < argc && std::string(argv[i]).rfind("--", 0) != 0) c.dir = argv[i++];
    for (; i < argc; ++i) {
        const std::string a = argv[i];
        auto need = [&](const char* what) {
            if (i + 1 >= argc) throw std::runtime_error(std::string(what) + " requires a value");
            return std::string(argv[++i]);
        };
        if (a == "--bitstream") c.bitstream = need("--bitstream");
        else if (a == "--tau") c.tau = std::stod(need("--tau"));
        else if (a == "--max-iters") c.max_iters = std::stoi(need("--max-iters"));
        else if (a == "--diff-tol") c.diff_tol = std::stod(need("--diff-tol"));
        else if (a == "--spmv-only") c.spmv_only = true;
        else if (a == "--pcg-spmv-service") c.pcg_spmv_service = true;
        else if (a == "--spmv-repeats") c.spmv_repeats = std::stoi(need("--spmv-repeats"));
        else throw std::runtime_error("unknown argument: " + a);
    }
    if (c.tau <= 0.0) throw std::runtime_error("--tau must be positive");
    if (c.max_iters < 0) throw std::runtime_error("--max-iters must be non-negative");
    if (c.diff_tol <= 0.0) throw std::runtime_error("--diff-tol must be positive");
    if (c.spmv_repeats <= 0) throw std::runtime_error("--spmv-repeats must be positive");
    return c;
}

void check(int rc) {
    if (rc != 0) {
        char msg[1024];
        xp_last_error(msg, sizeof msg);
        throw std::runtime_error(msg);
    }
}

void hip(hipError_t e) {
    if (e != hipSuccess) throw std::runtime_error(hipGetErrorString(e));
}

struct Diff {
    double abs = 0.0, rel = 0.0;
};

Diff max_diff(const std::vector<double>& got, const std::vector<double>& want) {
    Diff d;
    for (size_t k = 0; k < got.size(); ++k) {
        const double a = std::fabs(got[k] - want[k]);
        d.abs = std::max(d.abs, a);
        d.rel = std::max(d.rel, a / std::max(std::fabs(want[k]), 1.0e-12));
    }
    return d;
}

int spmv_only(const Cli& cli, const xp::CsrDataset& d) {
    const auto t0 = std::chrono::steady_clock::now();
    // plan: the reference's packed format and back, then the device matrix
    xp_cuper_packed* packed = nullptr;
    check(xp_cuper_pack(d.n(), d.row_ptr().data(), d.col_idx().data(), d.values().data(), &packed));
    xp_csr* imported = nullptr;
    check(xp_cuper_unpack(packed, d.n(), &imported));
    xp_cuper_packed_free(packed);
    int n = 0, nnz = 0;
    check(xp_csr_info(imported, &n, &nnz));
    std::vector<int> rp(static_cast<size_t>(n) + 1), ci(static_cast<size_t>(nnz));
    std::vector<double> va(static_cast<size_t>(nnz)), bz(static_cast<size_t>(n)), xz(static_cast<size_t>(n));
    check(xp_csr_copy(imported, rp.data(), ci.data(), va.data(), bz.data(), xz.data()));
    xp_csr_free(imported);
    xp_matrix* m = nullptr;
    xp_matrix_opts mo{};
    check(xp_matrix_create(n, rp.data(), ci.data(), va.data(), &mo, &m));
    xp_matrix_info info{};
    check(xp_matrix_get_info(m, &info));
    hip(hipDeviceSynchronize());
    const double plan_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();

    // input fixed to the dataset's b; the FP64 product as the check
    std::vector<double> expected;
    d.spmv(d.b(), expected);
    const size_t np = static_cast<size_t>(info.n_pad);
    void *bn = nullptr, *xs = nullptr, *ys = nullptr, *yn = nullptr;
    check(xp_device_alloc(static_cast<unsigned long long>(n) * 8, &bn));
    check(xp_device_alloc(np * 4, &xs));
    check(xp_device_alloc(np * 4, &ys));
    check(xp_device_alloc(static_cast<unsigned long long>(n) * 8, &yn));
    check(xp_memcpy_htod(bn, d.b().data(), static_cast<unsigned long long>(n) * 8));
    // every operation and both events on one stream of this program's own
    hipStream_t st = nullptr;
    hip(hipStreamCreate(&st));
    check(xp_to_solve_order(m, static_cast<const double*>(bn), static_cast<float*>(xs), st));
    hipEvent_t e0, e1;
    hip(hipEventCreate(&e0));
    hip(hipEventCreate(&e1));
    double spmv_ms = 0.0;
    for (int r = 0; r < cli.spmv_repeats; ++r) {
        hip(hipEventRecord(e0, st));
        check(xp_spmv_f32(m, static_cast<const float*>(xs), static_cast<float*>(ys), st));
        hip(hipEventRecord(e1, st));
        hip(hipEventSynchronize(e1));
        float ms = 0.0f;
        hip(hipEventElapsedTime(&ms, e0, e1));
        spmv_ms += ms;
    }
    check(xp_from_solve_order(m, static_cast<const float*>(ys), static_cast<double*>(yn), st));
    hip(hipStreamSynchronize(st));
    std::vector<double> output(static_cast<size_t>(n));
    check(xp_memcpy_dtoh(output.data(), yn, static_cast<unsigned long long>(n) * 8));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    for (void* p : {bn, xs, ys, yn}) xp_device_free(p);
    xp_matrix_destroy(m);

    const Diff df = max_diff(output, expected);
    const double kernel_seconds = spmv_ms * 1.0e-3;
    const double gflops =
        kernel_seconds > 0.0 ? 2.0 * static_cast<double>(d.nnz()) * cli.spmv_repeats / kernel_seconds / 1.0e9 : 0.0;
    std::cout << std::scientific << std::setprecision(12);
    std::cout << "[done] mode=spmv_only spmv_calls=" << std::defaultfloat << cli.spmv_repeats << std::scientific
              << " status=ok\n";
    std::cout << "[check] max_abs_diff=" << df.abs << " max_rel_diff=" << df.rel << " diff_tol=" << cli.diff_tol
              << "\n";
    std::cout << "[timing-ms] plan=" << plan_ms << " spmv_total=" << spmv_ms << " spmv_calls=" << std::defaultfloat
              << cli.spmv_repeats << std::scientific << " spmv_avg=" << spmv_ms / cli.spmv_repeats
              << " gflops=" << gflops << "\n";
    return (df.abs > cli.diff_tol && df.rel > cli.diff_tol) ? 3 : 0;
}

int device_pcg(const Cli& cli, const xp::CsrDataset& d) {
    const auto t0 = std::chrono::steady_clock::now();
    xp_matrix* m = nullptr;
    xp_matrix_opts mo{};
    mo.keep_lo = 1;  // hi + lo values: the true residual in FP64
    check(xp_matrix_create(d.n(), d.row_ptr().data(), d.col_idx().data(), d.values().data(), &mo, &m));
    xp_solve_opts so{};
    so.tau = cli.tau;
    so.max_iters = cli.max_iters;
    so.use_graph = 1;
    so.fused = 1;
    xp_pcg* s = nullptr;
    check(xp_pcg_create(m, &so, &s));
    hip(hipDeviceSynchronize());
    const double plan_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    check(xp_pcg_set_rhs(s, d.b().data(), d.x0().data()));
    check(xp_pcg_run(s));
    xp_status st{};
    check(xp_pcg_get_status(s, 0, &st));
    std::vector<double> x(static_cast<size_t>(d.n()));
    check(xp_pcg_get_x(s, x.data()));
    double res_abs = 0.0;
    check(xp_pcg_true_residual(s, &res_abs));
    const double bnorm = xp::norm2(d.b());
    // the SpMV-carrying kernel of the iteration, timed per launch
    double us[3] = {0.0, 0.0, 0.0};
    int nk = 0;
    check(xp_pcg_kernel_times(s, 20, us, &nk));
    xp_pcg_destroy(s);
    xp_matrix_destroy(m);

    xp::LoopSettings cfg;
    cfg.tau = cli.tau;
    cfg.max_iters = cli.max_iters;
    const xp::LoopResult golden = xp::run_host_pcg(d, cfg, nullptr, spmv_csr_kernel);
    const Diff df = max_diff(x, golden.x);
    const bool converged = st.status == XP_CONVERGED;
    const char* status = st.status == XP_CONVERGED ? "converged"
                         : st.status == XP_BREAKDOWN ? "breakdown" : "max_iter";
    const int calls = st.iterations + 1;
    std::cout << std::scientific << std::setprecision(12);
    std::cout << "[done] iter=" << st.iterations << " residual_abs=" << res_abs
              << " residual_rel=" << res_abs / std::max(bnorm, 1.0e-30) << " status=" << status << "\n";
    std::cout << "[check] cpu_residual_abs=" << golden.residual_l2 << " cuper_residual_abs=" << res_abs
              << " max_abs_diff=" << df.abs << " max_rel_diff=" << df.rel << " diff_tol=" << cli.diff_tol << "\n";
    std::cout << "[timing-ms] plan=" << plan_ms << " spmv_total=" << us[0] * 1.0e-3 * calls
              << " spmv_calls=" << std::defaultfloat << calls << std::scientific << " spmv_avg=" << us[0] * 1.0e-3
              << "\n";
    if (!converged) return 2;
    return (df.abs > cli.diff_tol && df.rel > cli.diff_tol) ? 3 : 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        const Cli cli = parse(argc, argv);
        const xp::CsrDataset d = xp::CsrDataset::load(cli.dir);
        std::cout << "[xplus] dataset=" << cli.dir << " mode=" << (cli.spmv_only ? "cuper-spmv-tapa" : "cuper-pcg-tapa")
                  << " spmv=" << (cli.pcg_spmv_service ? "tapa-cuper-compat-demo" : "tapa-cuper")
                  << " bitstream=" << (cli.bitstream.empty() ? "<none>" : cli.bitstream) << " device=mi355x\n";
        return cli.spmv_only ? spmv_only(cli, d) : device_pcg(cli, d);
    } catch (const std::exception& e) {
        usage(argv[0]);
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
}

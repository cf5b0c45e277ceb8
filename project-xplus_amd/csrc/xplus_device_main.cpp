// xplus_device_main.cpp -- command-line driver of the whole-solve device
// path, with the reference FPGA host's command line and reports
// (archived/host/xrt_host.cpp:60-175, 600-720):
//   [xclbin] [dataset_dir] [--tau v] [--max-iters v] [--device-index v]
//   [--timing] [--verbose] [--json-out path] [--txt-out path]
// The xclbin argument names the FPGA image there; it is accepted and
// recorded here.  One solve = one graph-captured PCG run on the device (the
// accurate iteration: FP64 x, p, A p and FP64-accumulated SpMV over hi + lo
// values, the counterpart of the reference's single control kernel with
// data_t = double), checked against the FP64 multi-kernel drop-in solver
// (the reference's golden arithmetic, the role of its CPU golden there).
// Same [done] / [check] / [host-timing-ms] / [kernel-timing-ms] lines, same
// reports (report_io.hpp), same exit codes: 0 pass, 1 error, 2 not
// converged, 3 max |x - x_golden| > 1e-8.
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
#include "report_io.hpp"

namespace {

namespace fs = std::filesystem;
using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Opts {
    fs::path xclbin = "mi355x";
    fs::path dataset = "data/suitesparse/Schmid/csr/thermal2_n1024";
    double tau = 1.0e-10;
    int max_iters = 0;
    unsigned device_index = 0;
    bool timing = true;
    bool verbose = false;
    fs::path json_out, txt_out;
};

double parse_double(const char* t, const char* name) {
    char* end = nullptr;
    const double v = std::strtod(t, &end);
    if (end == t || *end != '\0') throw std::runtime_error(std::string("invalid ") + name + ": " + t);
    return v;
}

long parse_long(const char* t, const char* name) {
    char* end = nullptr;
    const long v = std::strtol(t, &end, 10);
    if (end == t || *end != '\0') throw std::runtime_error(std::string("invalid ") + name + ": " + t);
    return v;
}

void usage(const char* argv0) {
    std::cerr << "Usage: " << argv0
              << " [xclbin] [dataset_dir] [--tau value] [--max-iters value] [--device-index value]"
              << " [--timing] [--verbose] [--json-out path] [--txt-out path]\n";
}

Opts parse(int argc, char** argv) {
    Opts o;
    int i = 1;
    if (i < argc && std::string(argv[i]).rfind("--", 0) != 0) o.xclbin = argv[i++];
    if (i < argc && std::string(argv[i]).rfind("--", 0) != 0) o.dataset = argv[i++];
    // reports/<prefix>.{json,txt} by default, as the reference's run defaults
    const std::string prefix = "MI355X_" + o.dataset.filename().string();
    o.json_out = fs::path("reports") / (prefix + ".json");
    o.txt_out = fs::path("reports") / (prefix + ".txt");
    for (; i < argc; ++i) {
        const std::string a = argv[i];
        auto need = [&](const char* what) {
            if (i + 1 >= argc) throw std::runtime_error(std::string(what) + " requires a value");
            return argv[++i];
        };
        if (a == "--tau") o.tau = parse_double(need("--tau"), "tau");
        else if (a == "--max-iters") o.max_iters = static_cast<int>(parse_long(need("--max-iters"), "max_iters"));
        else if (a == "--device-index") o.device_index = static_cast<unsigned>(parse_long(need("--device-index"), "device_index"));
        else if (a == "--timing") o.timing = true;
        else if (a == "--verbose") o.verbose = true;
        else if (a == "--json-out") o.json_out = need("--json-out");
        else if (a == "--txt-out") o.txt_out = need("--txt-out");
        else throw std::runtime_error("unknown argument: " + a);
    }
    if (o.tau <= 0.0) throw std::runtime_error("--tau must be positive");
    if (o.max_iters < 0) throw std::runtime_error("--max-iters must be non-negative");
    return o;
}

void check(int rc) {
    if (rc != 0) {
        char msg[1024];
        xp_last_error(msg, sizeof msg);
        throw std::runtime_error(msg);
    }
}

}  // namespace

int main(int argc, char** argv) {
    try {
        const Opts opt = parse(argc, argv);
        int devices = 0;
        if (hipGetDeviceCount(&devices) != hipSuccess || devices <= 0) throw std::runtime_error("no GPU visible");
        if (static_cast<int>(opt.device_index) >= devices)
            throw std::runtime_error("--device-index out of range: " + std::to_string(opt.device_index));
        if (hipSetDevice(static_cast<int>(opt.device_index)) != hipSuccess) throw std::runtime_error("hipSetDevice failed");

        const auto t_total = Clock::now();
        const xp::CsrDataset d = xp::CsrDataset::load(opt.dataset);
        std::cout << "[xplus] dataset=" << opt.dataset << " xclbin=" << opt.xclbin
                  << " mode=device-pcg device=mi355x\n";

        const auto t_setup = Clock::now();
        const int effective_max = opt.max_iters > 0 ? opt.max_iters : std::max(4 * d.n(), 1000);
        xp_matrix* m = nullptr;
        xp_matrix_opts mo{};
        mo.keep_lo = 1;
        check(xp_matrix_create(d.n(), d.row_ptr().data(), d.col_idx().data(), d.values().data(), &mo, &m));
        xp_solve_opts so{};
        so.tau = opt.tau;
        so.max_iters = effective_max;
        so.use_graph = 1;
        so.accurate = 2;
        so.trace_cap = std::min(effective_max, 1 << 20);   // 32 B per iteration on the device
        xp_pcg* s = nullptr;
        check(xp_pcg_create(m, &so, &s));
        check(xp_device_synchronize());
        const auto t_h2d = Clock::now();
        check(xp_pcg_set_rhs(s, d.b().data(), d.x0().data()));
        const auto t_kernel = Clock::now();
        check(xp_pcg_run(s));
        const auto t_d2h = Clock::now();
        std::vector<double> x(static_cast<size_t>(d.n()));
        check(xp_pcg_get_x(s, x.data()));
        xp_status st{};
        check(xp_pcg_get_status(s, 0, &st));
        std::vector<double> rows(static_cast<size_t>(std::max(st.iterations, 1)) * 4);
        const int got = xp_pcg_get_trace(s, 0, rows.data(), st.iterations);
        if (got < 0) check(got);
        const auto t_verify = Clock::now();

        xp::RunReport r;
        r.dataset_path = opt.dataset.string();
        r.xclbin_path = opt.xclbin.string();
        r.n = d.n();
        r.nnz = d.nnz();
        r.max_iters = effective_max;
        r.tau = opt.tau;
        r.device_index = opt.device_index;
        r.iterations = st.iterations;
        r.final_rr = st.rr;
        r.converged = st.status == XP_CONVERGED;
        r.breakdown = st.status == XP_BREAKDOWN;
        check(xp_pcg_true_residual(s, &r.residual_l2));
        r.residual_rel = r.residual_l2 / std::max(xp::norm2(d.b()), xp::kBreakdownEps);
        xp::LoopSettings cfg;
        cfg.tau = opt.tau;
        cfg.max_iters = opt.max_iters;
        const xp::LoopResult golden = xp::run_host_pcg(d, cfg, nullptr, spmv_csr_kernel);
        for (size_t k = 0; k < x.size(); ++k) r.max_abs_diff = std::max(r.max_abs_diff, std::fabs(x[k] - golden.x[k]));
        r.golden_residual_l2 = golden.residual_l2;
        r.pass = r.converged && r.max_abs_diff <= 1.0e-8;
        for (int k = 0; k < got; ++k) {
            const double* t = rows.data() + static_cast<size_t>(k) * 4;
            r.traces.push_back(xp::IterationTrace{k, t[0], t[1], t[2], t[3], std::sqrt(std::max(t[3], 0.0))});
        }
        const auto t_end = Clock::now();

        double solve_ms = 0.0;
        check(xp_pcg_last_timing(s, &solve_ms, nullptr));
        // per-kernel device times of the iteration (every launch timed with
        // events), scaled by the iteration count
        double us[3] = {0.0, 0.0, 0.0};
        int nk = 0;
        if (st.iterations > 0) check(xp_pcg_kernel_times(s, std::min(st.iterations, 50), us, &nk));
        xp_pcg_destroy(s);
        xp_matrix_destroy(m);

        r.total_ms = ms_since(t_total, t_end);
        r.host_setup_ms = ms_since(t_setup, t_h2d);
        r.h2d_ms = ms_since(t_h2d, t_kernel);
        r.kernel_ms = solve_ms;
        r.d2h_ms = ms_since(t_d2h, t_verify);
        r.verify_ms = ms_since(t_verify, t_end);
        r.kt.pcg_control_total_ms = solve_ms;
        r.kt.pcg_control_calls = 1;
        r.kt.spmv_calls = st.iterations;           // SpMV + pAp + direction update, one kernel
        r.kt.spmv_total_ms = us[0] * 1.0e-3 * st.iterations;
        r.kt.update_xrz_calls = st.iterations;     // x, r, z, rz, rr
        r.kt.update_xrz_total_ms = us[1] * 1.0e-3 * st.iterations;

        const char* status = r.breakdown ? "breakdown" : (r.converged ? "converged" : "max_iter");
        std::cout << std::scientific << std::setprecision(12);
        std::cout << "[done] iter=" << r.iterations << " residual_abs=" << r.residual_l2
                  << " residual_rel=" << r.residual_rel << " status=" << status << "\n";
        std::cout << "[check] cpu_residual_abs=" << r.golden_residual_l2 << " fpga_residual_abs=" << r.residual_l2
                  << " max_abs_diff=" << r.max_abs_diff << "\n";
        if (opt.timing) {
            std::cout << std::fixed << std::setprecision(3);
            std::cout << "[host-timing-ms] total=" << r.total_ms << " host_setup=" << r.host_setup_ms
                      << " buffer_h2d=" << r.h2d_ms << " kernel_total=" << r.kernel_ms << " buffer_d2h=" << r.d2h_ms
                      << " verify=" << r.verify_ms << "\n";
            const xp::KernelTimingStats& k = r.kt;
            auto avg = [](double t, int c) { return c > 0 ? t / c : 0.0; };
            std::cout << "[kernel-timing-ms] pcg_control_total=" << k.pcg_control_total_ms
                      << " pcg_control_avg=" << avg(k.pcg_control_total_ms, k.pcg_control_calls)
                      << " spmv_total=" << k.spmv_total_ms << " spmv_avg=" << avg(k.spmv_total_ms, k.spmv_calls)
                      << " init_total=" << k.init_total_ms << " dot_total=" << k.dot_total_ms
                      << " dot_avg=" << avg(k.dot_total_ms, k.dot_calls) << " update_xrz_total=" << k.update_xrz_total_ms
                      << " update_xrz_avg=" << avg(k.update_xrz_total_ms, k.update_xrz_calls)
                      << " update_p_total=" << k.update_p_total_ms
                      << " update_p_avg=" << avg(k.update_p_total_ms, k.update_p_calls) << "\n";
        }
        if (opt.verbose) std::cout << "[verbose] trace_rows=" << r.traces.size() << "\n";
        if (!opt.txt_out.empty()) xp::write_text_report(opt.txt_out, r);
        if (!opt.json_out.empty()) xp::write_json_report(opt.json_out, r);
        if (!r.converged) return 2;
        if (r.max_abs_diff > 1.0e-8) return 3;
        return 0;
    } catch (const std::exception& e) {
        usage(argv[0]);
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
}

// xplus_local_main.cpp -- command-line driver of the drop-in path.
//
// Same command line as the reference's local multi-kernel executable
// (archived/host/main.cpp:51-90): [dataset_dir] [--tau v] [--max-iters v]
// [--spmv csr|blocked], plus --device-solver, which also runs the
// graph-captured FP32 device solver on the same dataset and reports the
// difference.  Exit codes follow the reference (main.cpp:137-148): 0 ok,
// 1 error/usage, 2 not converged.
#include <cmath>
#include <filesystem>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>

#include "host_loop.hpp"

namespace {

struct Cli {
    std::filesystem::path dir = "data/suitesparse/Schmid/csr/thermal2_n1024";
    double tau = 1.0e-10;
    int max_iters = 0;
    bool blocked = false;
    bool device_solver = false;
};

Cli parse(int argc, char** argv) {
    Cli c;
    int i = 1;
    if (i < argc && std::string(argv[i]).rfind("--", 0) != 0) c.dir = argv[i++];
    for (; i < argc; ++i) {
        const std::string a = argv[i];
        auto need = [&](const char* what) {
            if (i + 1 >= argc) throw std::runtime_error(std::string(what) + " requires a value");
            return std::string(argv[++i]);
        };
        if (a == "--tau") c.tau = std::stod(need("--tau"));
        else if (a == "--max-iters") c.max_iters = std::stoi(need("--max-iters"));
        else if (a == "--spmv") {
            const std::string v = need("--spmv");
            if (v == "csr") c.blocked = false;
            else if (v == "blocked") c.blocked = true;
            else throw std::runtime_error("unknown --spmv value: " + v + " (expected csr or blocked)");
        } else if (a == "--device-solver") c.device_solver = true;
        else throw std::runtime_error("unknown argument: " + a);
    }
    if (c.tau <= 0.0) throw std::runtime_error("--tau must be positive");
    if (c.max_iters < 0) throw std::runtime_error("--max-iters must be non-negative");
    return c;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        const Cli cli = parse(argc, argv);
        const xp::CsrDataset d = xp::CsrDataset::load(cli.dir);
        xp::LoopSettings cfg;
        cfg.tau = cli.tau;
        cfg.max_iters = cli.max_iters;
        std::cout << "[xplus] dataset=" << cli.dir << " mode=local-multi-kernel spmv="
                  << (cli.blocked ? "blocked" : "csr") << " device=mi355x\n";
        const xp::LoopResult res = xp::run_host_pcg(
            d, cfg, &std::cout, cli.blocked ? spmv_blocked_kernel : spmv_csr_kernel);
        std::cout << std::scientific << std::setprecision(12);
        std::cout << "[done] iter=" << res.steps << " residual_abs=" << res.residual_l2
                  << " residual_rel=" << res.residual_rel << " status=" << xp::outcome_name(res.outcome) << "\n";
        if (cli.device_solver) {
            xp_matrix* m = nullptr;
            xp_matrix_opts mo{};
            mo.keep_lo = 1;
            if (xp_matrix_create(d.n(), d.row_ptr().data(), d.col_idx().data(), d.values().data(), &mo, &m) != 0)
                throw std::runtime_error("xp_matrix_create failed");
            xp_solve_opts so{};
            so.tau = cli.tau;
            so.max_iters = cli.max_iters;
            so.use_graph = 1;
            xp_pcg* s = nullptr;
            if (xp_pcg_create(m, &so, &s) != 0 || xp_pcg_set_rhs(s, d.b().data(), d.x0().data()) != 0 ||
                xp_pcg_run(s) != 0) {
                char msg[512];
                xp_last_error(msg, sizeof msg);
                throw std::runtime_error(msg);
            }
            xp_status st{};
            xp_pcg_get_status(s, 0, &st);
            std::vector<double> x(static_cast<size_t>(d.n()));
            xp_pcg_get_x(s, x.data());
            double res_norm = 0.0;
            xp_pcg_true_residual(s, &res_norm);
            double ms = 0.0;
            xp_pcg_last_timing(s, &ms, nullptr);
            double md = 0.0;
            for (size_t k = 0; k < x.size(); ++k) md = std::max(md, std::fabs(x[k] - res.x[k]));
            std::cout << "[device] iter=" << st.iterations << " status=" << st.status << " rr=" << st.rr
                      << " residual_abs=" << res_norm << " solve_ms=" << ms << " max_abs_diff_vs_multi_kernel=" << md
                      << "\n";
            xp_pcg_destroy(s);
            xp_matrix_destroy(m);
        }
        return res.converged ? 0 : 2;
    } catch (const std::exception& e) {
        std::cerr << "Usage: " << argv[0] << " [dataset_dir] [--tau value] [--max-iters value] [--spmv csr|blocked]"
                  << " [--device-solver]\n";
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
}

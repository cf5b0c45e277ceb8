// report_io.hpp -- text and JSON run reports.
//
// The reference writes one text and one JSON report per run of its FPGA
// host (archived/host/report_io.hpp:15-273, called from
// archived/host/xrt_host.cpp:662-709), and renders the JSON as HTML
// (scripts/render_report.py).  The sections, keys, order, number formats
// (%.12e for accuracy figures and traces, %.3f for milliseconds) and the
// status / pass wording are the same here, so the reference's renderer and
// any script reading those reports read these; the trace comes from the
// device solver's own per-iteration record (trace_metadata.source =
// "hardware") rather than from a CPU re-run.
#pragma once

#include <cmath>
#include <cstdio>
#include <filesystem>
#include <stdexcept>
#include <string>
#include <vector>

namespace xp {

struct IterationTrace {
    int iteration = 0;       // 0-based, as the reference's golden trace
    double alpha = 0.0;
    double beta = 0.0;
    double rz = 0.0;         // after the iteration
    double rr = 0.0;
    double residual = 0.0;   // sqrt(rr)
};

struct KernelTimingStats {
    double pcg_control_total_ms = 0.0;
    double spmv_total_ms = 0.0;
    double init_total_ms = 0.0;
    double dot_total_ms = 0.0;
    double update_xrz_total_ms = 0.0;
    double update_p_total_ms = 0.0;
    int pcg_control_calls = 0;
    int spmv_calls = 0;
    int init_calls = 0;
    int dot_calls = 0;
    int update_xrz_calls = 0;
    int update_p_calls = 0;
    bool split() const {
        return spmv_calls > 0 || init_calls > 0 || dot_calls > 0 || update_xrz_calls > 0 || update_p_calls > 0;
    }
};

struct RunReport {
    std::string dataset_path;
    std::string xclbin_path;     // the reference's FPGA image slot; the device name here
    int n = 0;
    int nnz = 0;
    int max_iters = 0;
    double tau = 0.0;
    unsigned device_index = 0;
    bool pass = false;
    bool converged = false;
    bool breakdown = false;
    int iterations = 0;
    double final_rr = 0.0;
    double residual_l2 = 0.0;
    double residual_rel = 0.0;
    double max_abs_diff = 0.0;
    double golden_residual_l2 = 0.0;
    double total_ms = 0.0, host_setup_ms = 0.0, h2d_ms = 0.0, kernel_ms = 0.0, d2h_ms = 0.0, verify_ms = 0.0;
    KernelTimingStats kt;
    std::vector<IterationTrace> traces;
    std::string trace_source = "hardware";
};

namespace report_detail {

inline std::string json_escape(const std::string& in) {
    std::string out;
    out.reserve(in.size() + 8);
    for (char ch : in) {
        if (ch == '\\') out += "\\\\";
        else if (ch == '"') out += "\\\"";
        else if (ch == '\n') out += "\\n";
        else if (ch == '\r') out += "\\r";
        else if (ch == '\t') out += "\\t";
        else if (static_cast<unsigned char>(ch) < 0x20) {
            char u[8];
            std::snprintf(u, sizeof u, "\\u%04x", static_cast<unsigned>(static_cast<unsigned char>(ch)));
            out += u;
        } else out += ch;
    }
    return out;
}

inline std::string sci(double v) {
    char b[64];
    std::snprintf(b, sizeof b, "%.12e", v);
    return b;
}

inline std::string fix(double v) {
    char b[64];
    std::snprintf(b, sizeof b, "%.3f", v);
    return b;
}

inline double avg(double total, int calls) { return calls > 0 ? total / calls : 0.0; }

inline const char* status_word(const RunReport& r) {
    return r.breakdown ? "breakdown" : (r.converged ? "converged" : "max_iter");
}

struct File {
    FILE* f;
    explicit File(const std::filesystem::path& p, const char* kind) {
        if (!p.empty() && p.has_parent_path()) std::filesystem::create_directories(p.parent_path());
        f = std::fopen(p.string().c_str(), "w");
        if (f == nullptr) throw std::runtime_error(std::string("failed to open ") + kind + " report: " + p.string());
    }
    ~File() { std::fclose(f); }
    void put(const std::string& s) { std::fputs(s.c_str(), f); }
};

}  // namespace report_detail

inline void write_text_report(const std::filesystem::path& path, const RunReport& r) {
    using namespace report_detail;
    File o(path, "text");
    const KernelTimingStats& k = r.kt;
    o.put("[Problem]\n");
    o.put("dataset            : \"" + r.dataset_path + "\"\n");
    o.put("n / nnz            : " + std::to_string(r.n) + " / " + std::to_string(r.nnz) + "\n");
    o.put("max_iters          : " + std::to_string(r.max_iters) + "\n");
    o.put("tau                : " + sci(r.tau) + "\n");
    o.put("device_index       : " + std::to_string(r.device_index) + "\n\n");
    o.put("[Accuracy]\n");
    o.put(std::string("pass               : ") + (r.pass ? "yes" : "no") + "\n");
    o.put(std::string("converged          : ") + (r.converged ? "yes" : "no") + "\n");
    o.put(std::string("status             : ") + status_word(r) + "\n");
    o.put("iterations         : " + std::to_string(r.iterations) + "\n");
    o.put("final_rr           : " + sci(r.final_rr) + "\n");
    o.put("final_residual     : " + sci(r.residual_l2) + "\n");
    o.put("residual_rel       : " + sci(r.residual_rel) + "\n");
    o.put("max_abs_diff       : " + sci(r.max_abs_diff) + "\n");
    o.put("golden_residual    : " + sci(r.golden_residual_l2) + "\n\n");
    o.put("[Host Timing ms]\n");
    o.put("total              : " + fix(r.total_ms) + "\n");
    o.put("host_setup         : " + fix(r.host_setup_ms) + "\n");
    o.put("buffer_h2d         : " + fix(r.h2d_ms) + "\n");
    o.put("kernel_total       : " + fix(r.kernel_ms) + "\n");
    o.put("buffer_d2h         : " + fix(r.d2h_ms) + "\n");
    o.put("verify             : " + fix(r.verify_ms) + "\n");
    o.put("\n[Kernel Timing ms]\n");
    o.put("pcg_control_total  : " + fix(k.pcg_control_total_ms) + "\n");
    o.put("pcg_control_avg    : " + fix(avg(k.pcg_control_total_ms, k.pcg_control_calls)) + "\n");
    o.put(std::string("split_timing_note  : ") +
          (k.split() ? "split kernel timings are available" : "not available for single pcg_control_kernel runs") +
          "\n");
    o.put("spmv_total         : " + fix(k.spmv_total_ms) + "\n");
    o.put("spmv_avg           : " + fix(avg(k.spmv_total_ms, k.spmv_calls)) + "\n");
    o.put("init_total         : " + fix(k.init_total_ms) + "\n");
    o.put("init_avg           : " + fix(avg(k.init_total_ms, k.init_calls)) + "\n");
    o.put("dot_total          : " + fix(k.dot_total_ms) + "\n");
    o.put("dot_avg            : " + fix(avg(k.dot_total_ms, k.dot_calls)) + "\n");
    o.put("update_xrz_total   : " + fix(k.update_xrz_total_ms) + "\n");
    o.put("update_xrz_avg     : " + fix(avg(k.update_xrz_total_ms, k.update_xrz_calls)) + "\n");
    o.put("update_p_total     : " + fix(k.update_p_total_ms) + "\n");
    o.put("update_p_avg       : " + fix(avg(k.update_p_total_ms, k.update_p_calls)) + "\n");
}

inline void write_json_report(const std::filesystem::path& path, const RunReport& r) {
    using namespace report_detail;
    File o(path, "json");
    const KernelTimingStats& k = r.kt;
    auto kv = [&](const char* key, const std::string& v, bool last = false) {
        o.put(std::string("    \"") + key + "\": " + v + (last ? "\n" : ",\n"));
    };
    o.put("{\n  \"dataset\": {\n");
    kv("path", "\"" + json_escape(r.dataset_path) + "\"");
    kv("xclbin", "\"" + json_escape(r.xclbin_path) + "\"");
    kv("n", std::to_string(r.n));
    kv("nnz", std::to_string(r.nnz));
    kv("max_iters", std::to_string(r.max_iters));
    kv("tau", sci(r.tau));
    kv("device_index", std::to_string(r.device_index), true);
    o.put("  },\n  \"result\": {\n");
    kv("pass", r.pass ? "true" : "false");
    kv("converged", r.converged ? "true" : "false");
    kv("status", std::string("\"") + status_word(r) + "\"");
    kv("iterations", std::to_string(r.iterations));
    kv("final_rr", sci(r.final_rr));
    kv("final_residual_norm", sci(r.residual_l2));
    kv("residual_rel", sci(r.residual_rel));
    kv("solution_max_abs_diff", sci(r.max_abs_diff));
    kv("golden_residual_norm", sci(r.golden_residual_l2), true);
    o.put("  },\n  \"host_timing_ms\": {\n");
    kv("total", fix(r.total_ms));
    kv("host_setup", fix(r.host_setup_ms));
    kv("buffer_h2d", fix(r.h2d_ms));
    kv("kernel_total", fix(r.kernel_ms));
    kv("buffer_d2h", fix(r.d2h_ms));
    kv("verify", fix(r.verify_ms), true);
    o.put("  },\n  \"kernel_timing_ms\": {\n");
    const struct {
        const char* name;
        double total;
        int calls;
    } rows[] = {{"pcg_control", k.pcg_control_total_ms, k.pcg_control_calls},
                {"spmv", k.spmv_total_ms, k.spmv_calls},
                {"init", k.init_total_ms, k.init_calls},
                {"dot", k.dot_total_ms, k.dot_calls},
                {"update_xrz", k.update_xrz_total_ms, k.update_xrz_calls},
                {"update_p", k.update_p_total_ms, k.update_p_calls}};
    for (const auto& row : rows) {
        kv((std::string(row.name) + "_total").c_str(), fix(row.total));
        kv((std::string(row.name) + "_avg").c_str(), fix(avg(row.total, row.calls)));
        kv((std::string(row.name) + "_calls").c_str(), std::to_string(row.calls));
    }
    kv("split_timing_available", k.split() ? "true" : "false");
    kv("timing_model", k.split() ? "\"split_kernels\"" : "\"single_control_kernel\"", true);
    o.put("  },\n  \"trace_metadata\": {\n");
    kv("source", "\"" + std::string(r.traces.empty() ? "none" : r.trace_source) + "\"");
    kv("note", r.traces.empty() ? "\"no per-iteration trace was collected\""
                                : "\"per-iteration alpha, beta, rz and rr recorded by the device solver\"",
       true);
    o.put("  },\n  \"iterations_trace\": [\n");
    for (size_t i = 0; i < r.traces.size(); ++i) {
        const IterationTrace& t = r.traces[i];
        o.put("    {\n");
        o.put("      \"iteration\": " + std::to_string(t.iteration) + ",\n");
        o.put("      \"alpha\": " + sci(t.alpha) + ",\n");
        o.put("      \"beta\": " + sci(t.beta) + ",\n");
        o.put("      \"rz\": " + sci(t.rz) + ",\n");
        o.put("      \"rr\": " + sci(t.rr) + ",\n");
        o.put("      \"residual\": " + sci(t.residual) + "\n");
        o.put(std::string("    }") + (i + 1 == r.traces.size() ? "\n" : ",\n"));
    }
    o.put("  ]\n}\n");
}

}  // namespace xp

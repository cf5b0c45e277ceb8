// ref_shim_cuper.cpp -- C entry point over the reference's software Cuper
// slice/window SpMV (host/cuper_pcg_solver.hpp:90-159), compiled by
// oracle/Makefile from the reference tree.  TEST INFRASTRUCTURE ONLY.
#include "cuper_pcg_solver.hpp"

#include <cstdio>
#include <cstring>
#include <exception>

namespace pc = project_xplus::cgsolver;

extern "C" int ref_cuper_style_spmv(const char* dir, const double* x, double* y, char* err,
                                    int errlen) {
    try {
        const pc::CsrDataset d = pc::CsrDataset::load(dir);
        const auto plan = pc::build_cuper_matrix_plan(d);
        std::vector<double> xv(x, x + d.n()), yv;
        pc::cuper_style_spmv(plan, xv, yv);
        std::memcpy(y, yv.data(), yv.size() * sizeof(double));
        return 0;
    } catch (const std::exception& e) {
        if (err != nullptr && errlen > 0) std::snprintf(err, static_cast<size_t>(errlen), "%s", e.what());
        return -1;
    }
}

// build_cuper_control_matrix (host/cuper_control_matrix.hpp:136-207), the
// reference's own packer, over a dataset directory: sizes first, then the
// arrays (sp_element_list_ptr and the 16 channels, channel-major).
#include "cuper_control_matrix.hpp"

extern "C" int ref_cuper_pack_info(const char* dir, int* batch_count, int* matrix_len, int* channel_size,
                                   char* err, int errlen) {
    try {
        const pc::CsrDataset d = pc::CsrDataset::load(dir);
        const auto m = pc::build_cuper_control_matrix(d);
        *batch_count = m.batch_count;
        *matrix_len = m.matrix_len;
        *channel_size = m.matrix_data.empty() ? 0 : static_cast<int>(m.matrix_data[0].size());
        return 0;
    } catch (const std::exception& e) {
        if (err != nullptr && errlen > 0) std::snprintf(err, static_cast<size_t>(errlen), "%s", e.what());
        return -1;
    }
}

extern "C" int ref_cuper_pack_copy(const char* dir, int* ptr, unsigned long long* data, char* err, int errlen) {
    try {
        const pc::CsrDataset d = pc::CsrDataset::load(dir);
        const auto m = pc::build_cuper_control_matrix(d);
        for (size_t i = 0; i < m.sp_element_list_ptr.size(); ++i) ptr[i] = static_cast<int>(m.sp_element_list_ptr[i]);
        size_t off = 0;
        for (const auto& ch : m.matrix_data) {
            for (size_t i = 0; i < ch.size(); ++i) data[off + i] = static_cast<unsigned long long>(ch[i]);
            off += ch.size();
        }
        return 0;
    } catch (const std::exception& e) {
        if (err != nullptr && errlen > 0) std::snprintf(err, static_cast<size_t>(errlen), "%s", e.what());
        return -1;
    }
}

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

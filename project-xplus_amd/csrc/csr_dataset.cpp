// csr_dataset.cpp -- loader/validator of the CSR text datasets.
#include "csr_dataset.hpp"

#include <cmath>
#include <cstdio>
#include <fstream>
#include <limits>
#include <stdexcept>

#include "../../include/xplus_capi.h"

namespace xp {

namespace fs = std::filesystem;

namespace {

template <typename T>
std::vector<T> read_array_file(const fs::path& path) {
    std::ifstream input(path);
    if (!input) throw std::runtime_error("failed to open " + path.string());
    std::vector<T> out;
    T value{};
    while (input >> value) out.push_back(value);
    if (out.empty()) throw std::runtime_error("empty array file: " + path.string());
    return out;
}

template <typename T>
void write_array_file(const fs::path& path, const std::vector<T>& v, bool is_float) {
    FILE* f = std::fopen(path.string().c_str(), "w");
    if (f == nullptr) throw std::runtime_error("failed to open " + path.string());
    for (size_t i = 0; i < v.size(); ++i) {
        if (i) std::fputc(' ', f);
        if (is_float)
            std::fprintf(f, "%.17g", static_cast<double>(v[i]));
        else
            std::fprintf(f, "%lld", static_cast<long long>(v[i]));
    }
    std::fputc('\n', f);
    std::fclose(f);
}

}  // namespace

void CsrDataset::validate_matrix() const {
    // Same checks, same order, same messages as src/CsrDataset.hpp:25-54.
    if (row_ptr_.size() < 2) throw std::runtime_error("row_ptr.txt must contain at least 2 entries");
    if (values_.size() != col_idx_.size()) throw std::runtime_error("col_idx.txt and values.txt size mismatch");
    if (row_ptr_.front() != 0) throw std::runtime_error("row_ptr[0] must be 0");
    if (row_ptr_.back() != nnz_) throw std::runtime_error("row_ptr[n] must equal nnz");
    for (int row = 0; row < n_; ++row) {
        const int rb = row_ptr_[static_cast<size_t>(row)];
        const int re = row_ptr_[static_cast<size_t>(row) + 1];
        if (rb > re) throw std::runtime_error("row_ptr must be nondecreasing");
        for (int o = rb; o < re; ++o) {
            const int c = col_idx_[static_cast<size_t>(o)];
            if (c < 0 || c >= n_) throw std::runtime_error("col_idx out of range");
        }
    }
}

void CsrDataset::validate_vectors() const {
    if (static_cast<int>(b_.size()) != n_) throw std::runtime_error("b.txt length must equal n");
    if (static_cast<int>(x0_.size()) != n_) throw std::runtime_error("x0.txt length must equal n");
}

CsrDataset CsrDataset::load(const fs::path& dir) {
    CsrDataset d;
    d.row_ptr_ = read_array_file<int>(dir / "row_ptr.txt");
    d.col_idx_ = read_array_file<int>(dir / "col_idx.txt");
    d.values_ = read_array_file<double>(dir / "values.txt");
    if (d.row_ptr_.size() < 2) throw std::runtime_error("row_ptr.txt must contain at least 2 entries");
    d.n_ = static_cast<int>(d.row_ptr_.size()) - 1;
    d.nnz_ = static_cast<int>(d.col_idx_.size());
    d.validate_matrix();
    d.b_ = read_array_file<double>(dir / "b.txt");
    d.x0_ = read_array_file<double>(dir / "x0.txt");
    d.validate_vectors();
    return d;
}

CsrDataset CsrDataset::from_arrays(std::vector<int> row_ptr, std::vector<int> col_idx,
                                   std::vector<double> values, std::vector<double> b,
                                   std::vector<double> x0) {
    CsrDataset d;
    d.row_ptr_ = std::move(row_ptr);
    d.col_idx_ = std::move(col_idx);
    d.values_ = std::move(values);
    d.b_ = std::move(b);
    d.x0_ = std::move(x0);
    if (d.row_ptr_.size() < 2) throw std::runtime_error("row_ptr.txt must contain at least 2 entries");
    d.n_ = static_cast<int>(d.row_ptr_.size()) - 1;
    d.nnz_ = static_cast<int>(d.col_idx_.size());
    d.validate_matrix();
    d.validate_vectors();
    return d;
}

void CsrDataset::save(const fs::path& dir) const {
    fs::create_directories(dir);
    write_array_file(dir / "row_ptr.txt", row_ptr_, false);
    write_array_file(dir / "col_idx.txt", col_idx_, false);
    write_array_file(dir / "values.txt", values_, true);
    write_array_file(dir / "b.txt", b_, true);
    write_array_file(dir / "x0.txt", x0_, true);
}

void CsrDataset::spmv(const std::vector<double>& x, std::vector<double>& y) const {
    y.assign(static_cast<size_t>(n_), 0.0);
    if (static_cast<int>(x.size()) < n_) throw std::runtime_error("spmv: x shorter than n");
    if (xp_csr_spmv_f64(n_, row_ptr_.data(), col_idx_.data(), values_.data(), x.data(), y.data()) != 0) {
        char msg[512];
        xp_last_error(msg, sizeof msg);
        throw std::runtime_error(std::string("spmv failed: ") + msg);
    }
}

std::vector<double> CsrDataset::extract_jacobi_diag() const {
    std::vector<double> diag(static_cast<size_t>(n_), 0.0);
    for (int row = 0; row < n_; ++row) {
        bool found = false;
        for (int o = row_ptr_[static_cast<size_t>(row)]; o < row_ptr_[static_cast<size_t>(row) + 1]; ++o) {
            if (col_idx_[static_cast<size_t>(o)] == row) {
                diag[static_cast<size_t>(row)] = values_[static_cast<size_t>(o)];
                found = true;
                break;
            }
        }
        if (!found) throw std::runtime_error("missing diagonal entry at row " + std::to_string(row));
        if (std::fabs(diag[static_cast<size_t>(row)]) <= std::numeric_limits<double>::min())
            throw std::runtime_error("zero diagonal entry at row " + std::to_string(row));
    }
    return diag;
}

}  // namespace xp

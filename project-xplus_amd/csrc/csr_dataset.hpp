// csr_dataset.hpp -- the dataset loader of the drop-in surface.
//
// Mirrors project_xplus::cgsolver::CsrDataset (src/CsrDataset.hpp:17-146):
// the same five text files (row_ptr.txt col_idx.txt values.txt b.txt
// x0.txt), the same whitespace parsing (`stream >> value` until the first
// token that does not parse), the same validation order and the same
// std::runtime_error messages, so a caller that catches them sees no change.
// spmv() runs on the GPU (FP64 SELL kernel); there is no CPU compute path.
#pragma once

#include <filesystem>
#include <string>
#include <vector>

namespace xp {

class CsrDataset {
  public:
    static CsrDataset load(const std::filesystem::path& dataset_dir);
    // Builds a dataset from arrays, with the same validation as load().
    static CsrDataset from_arrays(std::vector<int> row_ptr, std::vector<int> col_idx,
                                  std::vector<double> values, std::vector<double> b,
                                  std::vector<double> x0);

    int n() const { return n_; }
    int nnz() const { return nnz_; }
    const std::vector<int>& row_ptr() const { return row_ptr_; }
    const std::vector<int>& col_idx() const { return col_idx_; }
    const std::vector<double>& values() const { return values_; }
    const std::vector<double>& b() const { return b_; }
    const std::vector<double>& x0() const { return x0_; }

    // y = A x in FP64 on the GPU (src/CsrDataset.hpp:78-90).
    void spmv(const std::vector<double>& x, std::vector<double>& y) const;
    // First diagonal entry per row; throws "missing diagonal entry at row R"
    // or "zero diagonal entry at row R" (|d| <= DBL_MIN), src/CsrDataset.hpp:92-116.
    std::vector<double> extract_jacobi_diag() const;

    // Writes the five files with %.17g values (round-trips every double).
    void save(const std::filesystem::path& dataset_dir) const;

  private:
    void validate_matrix() const;
    void validate_vectors() const;
    int n_ = 0;
    int nnz_ = 0;
    std::vector<int> row_ptr_;
    std::vector<int> col_idx_;
    std::vector<double> values_;
    std::vector<double> b_;
    std::vector<double> x0_;
};

}  // namespace xp

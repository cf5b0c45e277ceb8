// mlp.hpp -- learned 8x8 block-Jacobi preconditioner (features, MLP, training).
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

#include "blocking.hpp"

namespace xp {

constexpr int kMlpIn = 48;      // 36 block entries + 8 off-block sums + 4 zero pad
constexpr int kMlpHidden = 64;
constexpr int kMlpOut = 48;     // 36 entries of the correction to L0 + 12 unused
constexpr double kMlpLambda = 0.5;  // weight of the off-block term of the training loss

// Features and training targets of the 8-row blocks of a matrix (rows in
// a block's slot order; padding slots are identity rows with scale 1).
struct BlockSet {
    int count = 0;
    std::vector<float> feat;   // count x kMlpIn
    std::vector<float> ahat;   // count x 64: A_bb / s_b
    std::vector<float> hmat;   // count x 64: Ahat^2 + lambda G
    std::vector<float> scale;  // count x 8: D^-1/2 of the block's rows (1 on padding rows)
    static BlockSet from_csr(int n, const int* row_ptr, const int* col_idx, const double* values);  // natural blocks
    static BlockSet from_csr(int n, const int* row_ptr, const int* col_idx, const double* values, const Blocking& b,
                             double lambda = kMlpLambda);
    void append(const BlockSet& o);
};

class Mlp {
  public:
    explicit Mlp(uint64_t seed);
    ~Mlp();
    static int num_params();
    const std::vector<float>& params();      // downloads the current weights
    void set_params(const float* p);
    // Mean loss over samples [first, first+count) and its gradient (flat, like params).
    double loss_grad(const BlockSet& data, int first, int count, float* grad);
    void forward(const BlockSet& data, int first, int count, float* out);  // count x kMlpOut
    std::vector<double> train(const BlockSet& data, int steps, int batch, float lr, uint64_t seed);
    // M_b = (L L^T + 1e-3 I) / s_b for every block of data (count x 64, row-major).
    std::vector<float> blocks(const BlockSet& data);
    // Device-side inference: feat [rows x kMlpIn] (rows = count rounded up to
    // 64, padding rows finite), scale [count x 8] -> blocks [count x 64].
    // All device pointers; synchronous.
    void blocks_device(const float* feat, const float* scale, int count, float* blocks);

  private:
    struct Dev;
    std::unique_ptr<Dev> d_;
    std::vector<float> params_;
    void upload_params();
    void cast_weights();
    void ensure_rows(int rows);
    int stage(const BlockSet& data, int first, int count);
    void forward_staged(int rows);
    void backward_staged(int rows, int count, const float* xf, const float* l0, const float* ahat, const float* hmat);
    double loss_grad_staged(int rows, int count);
};

}  // namespace xp

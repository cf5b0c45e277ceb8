// generators.hpp -- seeded, deterministic SPD test-matrix generators.
//
// The reference ships no PCG fixtures (its datasets are git-ignored
// downloads of SuiteSparse thermal2, data/suitesparse/SOURCES.md); the
// workloads are therefore generated here, bit-identically on any machine
// with this image: SplitMix64 streams and Box-Muller normals, no
// std::*_distribution.
#pragma once

#include <cstdint>
#include <vector>

namespace xp {

struct HostCsr {
    int n = 0;
    std::vector<int> row_ptr;
    std::vector<int> col_idx;
    std::vector<double> values;
    std::vector<double> b;
    std::vector<double> x0;
};

class Rng {
  public:
    explicit Rng(uint64_t seed) : s_(seed) {}
    uint64_t next() {
        uint64_t z = (s_ += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }   // [0, 1)
    double uniform(double lo, double hi) { return lo + (hi - lo) * uniform(); }
    double normal();  // Box-Muller, one variate per two uniforms
  private:
    uint64_t s_;
};

// Stream `k` of a seed, so that independent draws do not shift each other.
uint64_t stream_seed(uint64_t seed, uint64_t k);

// 2D 5-point variable-coefficient diffusion on an nx x ny grid, Dirichlet
// boundary through ghost edges on the diagonal, per-edge conductivity
// exp(sigma N(0,1)).  b ~ U[-1,1], x0 = 0.
HostCsr gen_diffusion2d(int nx, int ny, double sigma, uint64_t seed);

// 3D 7-point analogue on a k x k x k grid.
HostCsr gen_diffusion3d(int k, double sigma, uint64_t seed);

// P1 triangular FEM stiffness on an m x m node grid of the unit square:
// interior nodes jittered by U[-jitter h, jitter h] per coordinate, every
// cell split along a random diagonal, per-element conductivity
// exp(sigma N(0,1)) times `contrast` inside n_incl random discs, Dirichlet
// boundary nodes eliminated (n = (m-2)^2).  b ~ U[-1,1], x0 = 0.
HostCsr gen_fem2d(int m, double sigma, int n_incl, double contrast, double jitter, uint64_t seed);

}  // namespace xp

// blocking.cpp -- natural and strength-aggregated 8-row blocks (host C++).
#include "blocking.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace xp {

Blocking natural_blocking(int n) {
    Blocking b;
    b.n = n;
    const int nu = (n + 7) / 8;
    b.slots.assign(static_cast<size_t>(nu) * 8, -1);
    for (int r = 0; r < n; ++r) b.slots[static_cast<size_t>(r)] = r;
    return b;
}

namespace {

// Grows blocks greedily over the rows with eligible[r] (and not yet
// assigned), seeds in row order.  Appends the blocks to `out`.
void grow(int n, const int* row_ptr, const int* col_idx, const std::vector<double>& strength,
          const std::vector<char>& eligible, std::vector<int>& owner, std::vector<std::vector<int>>& out) {
    std::vector<std::pair<int, double>> cand;  // (row, summed strength to the block)
    for (int seed = 0; seed < n; ++seed) {
        if (!eligible[static_cast<size_t>(seed)] || owner[static_cast<size_t>(seed)] >= 0) continue;
        const int id = static_cast<int>(out.size());
        std::vector<int> blk = {seed};
        owner[static_cast<size_t>(seed)] = id;
        cand.clear();
        auto add_neighbours = [&](int v) {
            for (int o = row_ptr[v]; o < row_ptr[v + 1]; ++o) {
                const int j = col_idx[o];
                if (j == v || !eligible[static_cast<size_t>(j)] || owner[static_cast<size_t>(j)] >= 0) continue;
                const double w = strength[static_cast<size_t>(o)];
                bool found = false;
                for (auto& c : cand)
                    if (c.first == j) {
                        c.second += w;
                        found = true;
                        break;
                    }
                if (!found) cand.emplace_back(j, w);
            }
        };
        add_neighbours(seed);
        while (blk.size() < 8) {
            int best = -1;
            double bw = -1.0;
            size_t bi = 0;
            for (size_t i = 0; i < cand.size(); ++i) {
                const auto& c = cand[i];
                if (owner[static_cast<size_t>(c.first)] >= 0) continue;
                if (c.second > bw || (c.second == bw && c.first < best)) {
                    best = c.first;
                    bw = c.second;
                    bi = i;
                }
            }
            if (best < 0) break;
            cand.erase(cand.begin() + static_cast<std::ptrdiff_t>(bi));
            owner[static_cast<size_t>(best)] = id;
            blk.push_back(best);
            add_neighbours(best);
        }
        out.push_back(std::move(blk));
    }
}

}  // namespace

Blocking strength_blocking(int n, const int* row_ptr, const int* col_idx, const double* values) {
    if (n < 0) throw std::runtime_error("strength_blocking: negative n");
    Blocking b;
    b.n = n;
    if (n == 0) return b;
    if (row_ptr == nullptr || col_idx == nullptr || values == nullptr)
        throw std::runtime_error("strength_blocking: null CSR array");
    for (int r = 0; r < n; ++r)
        for (int o = row_ptr[r]; o < row_ptr[r + 1]; ++o)
            if (col_idx[o] < 0 || col_idx[o] >= n) throw std::runtime_error("col_idx out of range");
    // D^-1/2 from the first diagonal entry (1 where it is missing or not positive)
    std::vector<double> dis(static_cast<size_t>(n), 1.0);
    for (int r = 0; r < n; ++r)
        for (int o = row_ptr[r]; o < row_ptr[r + 1]; ++o)
            if (col_idx[o] == r) {
                if (values[o] > 0.0) dis[static_cast<size_t>(r)] = 1.0 / std::sqrt(values[o]);
                break;
            }
    std::vector<double> strength(static_cast<size_t>(row_ptr[n]));
    for (int r = 0; r < n; ++r)
        for (int o = row_ptr[r]; o < row_ptr[r + 1]; ++o)
            strength[static_cast<size_t>(o)] =
                std::fabs(values[o]) * dis[static_cast<size_t>(r)] * dis[static_cast<size_t>(col_idx[o])];

    std::vector<int> owner(static_cast<size_t>(n), -1);
    std::vector<std::vector<int>> first;
    grow(n, row_ptr, col_idx, strength, std::vector<char>(static_cast<size_t>(n), 1), owner, first);

    // pool the rows of partial blocks and regroup them among themselves
    std::vector<std::vector<int>> full, partial;
    std::vector<char> pool(static_cast<size_t>(n), 0);
    for (auto& blk : first) {
        if (blk.size() == 8) {
            full.push_back(std::move(blk));
        } else {
            for (int r : blk) {
                pool[static_cast<size_t>(r)] = 1;
                owner[static_cast<size_t>(r)] = -1;
            }
        }
    }
    std::vector<std::vector<int>> second;
    grow(n, row_ptr, col_idx, strength, pool, owner, second);
    for (auto& blk : second) {
        if (blk.size() == 8)
            full.push_back(std::move(blk));
        else
            partial.push_back(std::move(blk));
    }
    // first-fit packing of the rest, in creation order
    std::vector<std::vector<int>> packed;
    for (auto& blk : partial) {
        bool placed = false;
        for (auto& p : packed)
            if (p.size() + blk.size() <= 8) {
                p.insert(p.end(), blk.begin(), blk.end());
                placed = true;
                break;
            }
        if (!placed) packed.push_back(blk);
    }
    for (auto& p : packed) full.push_back(std::move(p));

    for (auto& blk : full) std::sort(blk.begin(), blk.end());
    std::sort(full.begin(), full.end(), [](const std::vector<int>& x, const std::vector<int>& y) { return x[0] < y[0]; });
    b.slots.assign(full.size() * 8, -1);
    for (size_t u = 0; u < full.size(); ++u)
        for (size_t j = 0; j < full[u].size(); ++j) b.slots[u * 8 + j] = full[u][j];
    return b;
}

}  // namespace xp

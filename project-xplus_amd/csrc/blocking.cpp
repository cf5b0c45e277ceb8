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
// assigned), seeds in the order of `seeds`.  With `total` non-null a
// candidate's score is its summed strength to the block divided by its
// total strength (the fraction of its coupling the block would hold),
// else the summed strength itself.  Appends the blocks to `out`.
void grow(const std::vector<int>& seeds, const int* row_ptr, const int* col_idx, const std::vector<double>& strength,
          const double* total, const std::vector<char>& eligible, std::vector<int>& owner,
          std::vector<std::vector<int>>& out) {
    std::vector<std::pair<int, double>> cand;  // (row, summed strength to the block)
    for (int seed : seeds) {
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
                double sc = c.second;
                if (total != nullptr) {
                    const double t = total[c.first];
                    sc = t > 0.0 ? c.second / t : 0.0;
                }
                if (sc > bw || (sc == bw && c.first < best)) {
                    best = c.first;
                    bw = sc;
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

Blocking strength_blocking(int n, const int* row_ptr, const int* col_idx, const double* values, int rule) {
    if (rule != kRowOrder && rule != kStrengthOrder) throw std::runtime_error("strength_blocking: unknown rule");
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

    std::vector<int> seeds(static_cast<size_t>(n));
    for (int r = 0; r < n; ++r) seeds[static_cast<size_t>(r)] = r;
    std::vector<double> total;
    if (rule == kStrengthOrder) {
        // total off-diagonal strength per row, summed in CSR order; seeds by
        // decreasing total (ties: the smaller row)
        total.assign(static_cast<size_t>(n), 0.0);
        for (int r = 0; r < n; ++r)
            for (int o = row_ptr[r]; o < row_ptr[r + 1]; ++o)
                if (col_idx[o] != r) total[static_cast<size_t>(r)] += strength[static_cast<size_t>(o)];
        std::stable_sort(seeds.begin(), seeds.end(), [&](int x, int y) {
            return total[static_cast<size_t>(x)] > total[static_cast<size_t>(y)];
        });
    }
    const double* tot = rule == kStrengthOrder ? total.data() : nullptr;

    std::vector<int> owner(static_cast<size_t>(n), -1);
    std::vector<std::vector<int>> first;
    grow(seeds, row_ptr, col_idx, strength, tot, std::vector<char>(static_cast<size_t>(n), 1), owner, first);

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
    grow(seeds, row_ptr, col_idx, strength, tot, pool, owner, second);
    for (auto& blk : second) {
        if (blk.size() == 8)
            full.push_back(std::move(blk));
        else
            partial.push_back(std::move(blk));
    }
    std::vector<std::vector<int>> packed;
    if (rule == kRowOrder) {
        // first-fit packing of the rest, in creation order
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
    } else {
        // pieces in order of their smallest row, each into the most recent of
        // the last kPackWindow packed blocks with room: merged pieces stay
        // close in row index (the seeds of this rule are scattered over the
        // matrix, so creation order is not local)
        constexpr size_t kPackWindow = 16;
        for (auto& blk : partial) std::sort(blk.begin(), blk.end());
        std::stable_sort(partial.begin(), partial.end(),
                         [](const std::vector<int>& x, const std::vector<int>& y) { return x[0] < y[0]; });
        for (auto& blk : partial) {
            bool placed = false;
            const size_t lo = packed.size() > kPackWindow ? packed.size() - kPackWindow : 0;
            for (size_t k = packed.size(); k-- > lo;)
                if (packed[k].size() + blk.size() <= 8) {
                    packed[k].insert(packed[k].end(), blk.begin(), blk.end());
                    placed = true;
                    break;
                }
            if (!placed) packed.push_back(blk);
        }
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

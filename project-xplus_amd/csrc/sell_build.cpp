// sell_build.cpp -- CSR -> SELL-C-sigma builder (host C++).
#include "sell_build.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>

namespace xp {

namespace {

int round_up(int v, int m) { return (v + m - 1) / m * m; }

}  // namespace

int16_t lo16_of(double v, float hi) {
    if (hi == 0.0f || !std::isfinite(hi)) return 0;
    int e = 0;
    (void)std::frexp(static_cast<double>(hi), &e);  // hi = f 2^e, 0.5 <= |f| < 1: E = e - 1
    const double q = std::ldexp(v - static_cast<double>(hi), 38 - (e - 1));
    const double r = std::nearbyint(q);
    return static_cast<int16_t>(std::max(-32767.0, std::min(32767.0, r)));
}

SellHost build_sell(const std::vector<CsrRef>& systems, int sigma, bool keep_lo, bool keep64, int block_rows,
                    const std::vector<Blocking>* blocks) {
    if (systems.empty()) throw std::runtime_error("build_sell: no systems");
    if (sigma <= 0) sigma = kC;
    if (sigma % kC != 0) throw std::runtime_error("build_sell: sigma must be a multiple of 64");
    if (block_rows != 1 && block_rows != 8) throw std::runtime_error("build_sell: block_rows must be 1 or 8");
    const int br = block_rows;

    if (blocks != nullptr && (br != 8 || blocks->size() != systems.size()))
        throw std::runtime_error("build_sell: blocks need block_rows = 8 and one Blocking per system");

    SellHost s;
    s.nsys = static_cast<int>(systems.size());
    s.sigma = sigma;
    s.block_rows = br;
    s.sys_row_begin.assign(systems.size() + 1, 0);
    s.sys_nat_begin.assign(systems.size() + 1, 0);
    s.unit_first.assign(systems.size() + 1, 0);
    // the units (blocks of br solve rows) of every system
    std::vector<Blocking> units(systems.size());
    for (size_t k = 0; k < systems.size(); ++k) {
        const CsrRef& a = systems[k];
        if (a.n < 0) throw std::runtime_error("build_sell: negative n");
        if (a.n > 0 && (a.row_ptr == nullptr || a.col_idx == nullptr || a.values == nullptr))
            throw std::runtime_error("build_sell: null CSR array");
        if (br == 8) {
            units[k] = blocks != nullptr ? (*blocks)[k] : natural_blocking(a.n);
            const Blocking& b = units[k];
            if (b.n != a.n || b.slots.size() % 8 != 0) throw std::runtime_error("build_sell: blocking does not match the system");
            std::vector<char> seen(static_cast<size_t>(a.n), 0);
            for (int r : b.slots) {
                if (r < -1 || r >= a.n || (r >= 0 && seen[static_cast<size_t>(r)]))
                    throw std::runtime_error("build_sell: blocking is not a partition of the rows");
                if (r >= 0) seen[static_cast<size_t>(r)] = 1;
            }
            for (char c : seen)
                if (!c) throw std::runtime_error("build_sell: blocking is not a partition of the rows");
        }
        const int nu = br == 8 ? units[k].units() : a.n;
        s.unit_first[k + 1] = s.unit_first[k] + (br == 8 ? nu : 0);
        // Every system gets at least one slice so that it owns a workgroup;
        // every block keeps its own 8 solve rows.
        const int padded = std::max(round_up(nu * br, kC), kC);
        s.sys_row_begin[k + 1] = s.sys_row_begin[k] + padded;
        s.sys_nat_begin[k + 1] = s.sys_nat_begin[k] + a.n;
    }
    s.n = s.sys_nat_begin.back();
    s.n_pad = s.sys_row_begin.back();
    s.nslices = s.n_pad / kC;
    s.perm.assign(static_cast<size_t>(s.n_pad), -1);
    s.iperm.assign(static_cast<size_t>(s.n), -1);
    s.has_diag.assign(static_cast<size_t>(s.n_pad), 1);

    // Off-diagonal count per natural row (every entry except the first
    // diagonal one), validated against the CSR invariants.
    std::vector<int> offcnt(static_cast<size_t>(s.n), 0);
    std::vector<int> diag_pos(static_cast<size_t>(s.n), -1);
    for (int k = 0; k < s.nsys; ++k) {
        const CsrRef& a = systems[static_cast<size_t>(k)];
        if (a.n == 0) continue;
        if (a.row_ptr[0] != 0) throw std::runtime_error("row_ptr[0] must be 0");
        for (int r = 0; r < a.n; ++r) {
            const int rb = a.row_ptr[r], re = a.row_ptr[r + 1];
            if (rb > re) throw std::runtime_error("row_ptr must be nondecreasing");
            const int g = s.sys_nat_begin[static_cast<size_t>(k)] + r;
            int cnt = 0;
            for (int o = rb; o < re; ++o) {
                const int c = a.col_idx[o];
                if (c < 0 || c >= a.n) throw std::runtime_error("col_idx out of range");
                if (c == r && diag_pos[static_cast<size_t>(g)] < 0) {
                    diag_pos[static_cast<size_t>(g)] = o;
                } else {
                    ++cnt;
                }
            }
            offcnt[static_cast<size_t>(g)] = cnt;
            s.nnz += re - rb;
        }
    }

    // Permutation: stable sort by descending off-diagonal count inside each
    // window of sigma rows of each system (units of br rows, keyed by the
    // longest row of the unit, when block_rows = 8).
    if (br == 8) s.block_unit.assign(static_cast<size_t>(s.n_pad / 8), -1);
    for (int k = 0; k < s.nsys; ++k) {
        const int nat0 = s.sys_nat_begin[static_cast<size_t>(k)];
        const int n_k = s.sys_nat_begin[static_cast<size_t>(k) + 1] - nat0;
        const int solve0 = s.sys_row_begin[static_cast<size_t>(k)];
        const int nu = br == 8 ? units[static_cast<size_t>(k)].units() : n_k;
        // local row of slot j of unit u (-1: padding)
        auto slot = [&](int u, int j) { return br == 8 ? units[static_cast<size_t>(k)].slots[static_cast<size_t>(u) * 8 + static_cast<size_t>(j)] : u; };
        std::vector<int> key(static_cast<size_t>(nu), 0);
        for (int u = 0; u < nu; ++u)
            for (int j = 0; j < br; ++j) {
                const int r = slot(u, j);
                if (r >= 0) key[static_cast<size_t>(u)] = std::max(key[static_cast<size_t>(u)], offcnt[static_cast<size_t>(nat0 + r)]);
            }
        std::vector<int> order(static_cast<size_t>(nu));
        std::iota(order.begin(), order.end(), 0);
        const int wu = sigma / br;  // units per window
        for (int w0 = 0; w0 < nu; w0 += wu) {
            const int w1 = std::min(nu, w0 + wu);
            std::stable_sort(order.begin() + w0, order.begin() + w1,
                             [&](int a, int b) { return key[static_cast<size_t>(a)] > key[static_cast<size_t>(b)]; });
        }
        for (int pos = 0; pos < nu; ++pos) {
            const int u = order[static_cast<size_t>(pos)];
            if (br == 8) s.block_unit[static_cast<size_t>(solve0 / 8 + pos)] = s.unit_first[static_cast<size_t>(k)] + u;
            for (int j = 0; j < br; ++j) {
                const int r = slot(u, j);
                if (r < 0) continue;
                const int g = nat0 + r;
                const int i = solve0 + pos * br + j;
                s.perm[static_cast<size_t>(i)] = g;
                s.iperm[static_cast<size_t>(g)] = i;
            }
        }
    }

    // Slice widths (1 diagonal column + max off-diagonal count).
    s.slice_ptr.assign(static_cast<size_t>(s.nslices) + 1, 0);
    for (int sl = 0; sl < s.nslices; ++sl) {
        int w = 0;
        for (int l = 0; l < kC; ++l) {
            const int g = s.perm[static_cast<size_t>(sl * kC + l)];
            if (g >= 0) w = std::max(w, offcnt[static_cast<size_t>(g)]);
        }
        s.slice_ptr[static_cast<size_t>(sl) + 1] = s.slice_ptr[static_cast<size_t>(sl)] + 1 + w;
    }
    s.stored = static_cast<long long>(s.slice_ptr.back()) * kC;
    const size_t total = static_cast<size_t>(s.stored);
    s.val.assign(total, 0.0f);
    if (keep_lo) s.val_lo.assign(total, 0.0f);
    if (keep_lo) s.val_lo16.assign(total, 0);
    if (keep64) s.val64.assign(total, 0.0);
    std::vector<int> cols(total, 0);
    s.slice_base.assign(static_cast<size_t>(s.nslices), 0);

    // System index of a global natural row, for column translation.
    std::vector<int> row_sys(static_cast<size_t>(s.n), 0);
    for (int k = 0; k < s.nsys; ++k)
        for (int g = s.sys_nat_begin[static_cast<size_t>(k)]; g < s.sys_nat_begin[static_cast<size_t>(k) + 1]; ++g)
            row_sys[static_cast<size_t>(g)] = k;

    bool fits16 = true;
    for (int sl = 0; sl < s.nslices; ++sl) {
        const size_t base = static_cast<size_t>(s.slice_ptr[static_cast<size_t>(sl)]) * kC;
        // The span always covers the slice's own rows: padding entries point
        // at their own row (value 0), so a non-finite x elsewhere cannot leak in.
        int cmin = sl * kC, cmax = sl * kC + kC - 1;
        for (int l = 0; l < kC; ++l) {
            const int i = sl * kC + l;  // solve row
            const int g = s.perm[static_cast<size_t>(i)];
            const size_t e0 = base + static_cast<size_t>(l);
            cols[e0] = i;
            if (g < 0) {  // padding row: identity
                s.val[e0] = 1.0f;
                if (keep64) s.val64[e0] = 1.0;
                continue;
            }
            const int k = row_sys[static_cast<size_t>(g)];
            const CsrRef& a = systems[static_cast<size_t>(k)];
            const int r = g - s.sys_nat_begin[static_cast<size_t>(k)];
            const int nat0 = s.sys_nat_begin[static_cast<size_t>(k)];
            const int dp = diag_pos[static_cast<size_t>(g)];
            const double dv = dp >= 0 ? a.values[dp] : 0.0;
            s.has_diag[static_cast<size_t>(i)] = dp >= 0 ? 1 : 0;
            s.val[e0] = static_cast<float>(dv);
            if (keep_lo) s.val_lo[e0] = static_cast<float>(dv - static_cast<double>(s.val[e0]));
            if (keep_lo) s.val_lo16[e0] = lo16_of(dv, s.val[e0]);
            if (keep64) s.val64[e0] = dv;
            int j = 1;
            for (int o = a.row_ptr[r]; o < a.row_ptr[r + 1]; ++o) {
                if (o == dp) continue;
                const size_t e = base + static_cast<size_t>(j) * kC + static_cast<size_t>(l);
                const int c = s.iperm[static_cast<size_t>(nat0 + a.col_idx[o])];
                const double v = a.values[o];
                s.val[e] = static_cast<float>(v);
                if (keep_lo) s.val_lo[e] = static_cast<float>(v - static_cast<double>(s.val[e]));
                if (keep_lo) s.val_lo16[e] = lo16_of(v, s.val[e]);
                if (keep64) s.val64[e] = v;
                cols[e] = c;
                cmin = std::min(cmin, c);
                cmax = std::max(cmax, c);
                ++j;
            }
        }
        s.slice_base[static_cast<size_t>(sl)] = cmin;
        const int w = s.slice_ptr[static_cast<size_t>(sl) + 1] - s.slice_ptr[static_cast<size_t>(sl)];
        for (int j = 1; j < w; ++j)
            for (int l = 0; l < kC; ++l) {
                const size_t e = base + static_cast<size_t>(j) * kC + static_cast<size_t>(l);
                const int g = s.perm[static_cast<size_t>(sl * kC + l)];
                const int used = g >= 0 ? offcnt[static_cast<size_t>(g)] : 0;
                if (j > used) cols[e] = sl * kC + l;
            }
        if (cmax - cmin > 65535) fits16 = false;
    }

    s.idx16 = fits16 ? 1 : 0;
    if (fits16) {
        s.col16.assign(total, 0);
        for (int sl = 0; sl < s.nslices; ++sl) {
            const size_t base = static_cast<size_t>(s.slice_ptr[static_cast<size_t>(sl)]) * kC;
            const size_t end = static_cast<size_t>(s.slice_ptr[static_cast<size_t>(sl) + 1]) * kC;
            const int cb = s.slice_base[static_cast<size_t>(sl)];
            for (size_t e = base + kC; e < end; ++e)
                s.col16[e] = static_cast<uint16_t>(cols[e] - cb);
        }
    } else {
        s.col32 = std::move(cols);
    }
    return s;
}

}  // namespace xp

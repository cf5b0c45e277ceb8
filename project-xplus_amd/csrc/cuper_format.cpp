// cuper_format.cpp -- Cuper packed matrix format: pack / unpack.
#include "cuper_format.hpp"

#include <algorithm>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>
#include <thread>

namespace xp {

namespace {

struct Elem {
    int col = -1;
    int row = -1;
    float value = 0.0f;
};

uint64_t pack_word(const Elem& e) {
    if (e.row < 0 || e.col < 0) return 0x3FFFFull << 32;
    uint32_t bits;
    std::memcpy(&bits, &e.value, 4);
    return (static_cast<uint64_t>(e.col) & 0x3FFFull) << 50 | (static_cast<uint64_t>(e.row) & 0x3FFFFull) << 32 |
           static_cast<uint64_t>(bits);
}

// One (batch, PE) list: column sort, then each entry goes to the first free
// slot at least kCuperWindow after the last slot of its row group
// (append_reordered_cuper_elements, host/cuper_control_matrix.hpp:82-127).
// window[] is indexed by row group and holds -kCuperWindow for groups not
// seen in this list; touched groups are reset before returning.
void schedule(std::vector<Elem>& in, int base_col, std::vector<int>& window, std::vector<int>& touched,
              std::vector<Elem>& out) {
    std::sort(in.begin(), in.end(), [](const Elem& a, const Elem& b) { return a.col < b.col; });
    out.clear();
    for (const Elem& e : in) {
        const int group = e.row / (2 * kCuperPes);
        int& last = window[static_cast<size_t>(group)];
        if (last == -kCuperWindow) touched.push_back(group);
        int slot = last + kCuperWindow;
        while (true) {
            if (slot >= static_cast<int>(out.size())) out.resize(static_cast<size_t>(slot) + 1);
            const Elem& cur = out[static_cast<size_t>(slot)];
            if (cur.row < 0 && cur.col < 0) break;
            ++slot;
        }
        out[static_cast<size_t>(slot)] = Elem{e.col - base_col, group * 2 + (e.row & 1), e.value};
        last = slot;
    }
    for (int g : touched) window[static_cast<size_t>(g)] = -kCuperWindow;
    touched.clear();
}

// f(0) .. f(workers - 1), one thread each; exceptions rethrown after the
// join, a thread that cannot be started runs its share inline.
template <typename F>
void for_workers(int workers, F&& f) {
    std::vector<std::exception_ptr> err(static_cast<size_t>(std::max(workers, 1)));
    auto run = [&](int w) {
        try {
            f(w);
        } catch (...) {
            err[static_cast<size_t>(w)] = std::current_exception();
        }
    };
    std::vector<std::thread> th;
    for (int w = 1; w < workers; ++w) {
        try {
            th.emplace_back(run, w);
        } catch (...) {
            run(w);
        }
    }
    run(0);
    for (auto& t : th) t.join();
    for (auto& e : err)
        if (e) std::rethrow_exception(e);
}

}  // namespace

int cuper_pe_for_row(int row) {
    // host/cuper_control_matrix.hpp:48-54
    const int packet = row / 2;
    const int checker = packet % 8;
    const int acc = (packet / 8) % 2;
    const int pe_in_acc = (packet / 16) % 8;
    return (checker * 2 + acc) * 8 + pe_in_acc;
}

CuperPacked cuper_pack(int n, const int* row_ptr, const int* col_idx, const double* values) {
    if (n < 0) throw std::runtime_error("n must be non-negative");
    if (n > kCuperMaxRows) throw std::runtime_error("matrix too large for the Cuper format (n > 16777216)");
    CuperPacked p;
    const int batches = (n + kCuperSliceWidth - 1) / kCuperSliceWidth;
    p.batch_count = batches;
    p.ptr.assign(static_cast<size_t>(batches) + 1, 0);
    // entries per (batch, PE) in CSR order
    std::vector<std::vector<Elem>> bucket(static_cast<size_t>(batches) * kCuperPes);
    for (int r = 0; r < n; ++r) {
        const int pe = cuper_pe_for_row(r);
        for (int o = row_ptr[r]; o < row_ptr[r + 1]; ++o) {
            const int c = col_idx[o];
            if (c < 0 || c >= n) throw std::runtime_error("col_idx out of range");
            bucket[static_cast<size_t>(c / kCuperSliceWidth) * kCuperPes + static_cast<size_t>(pe)].push_back(
                Elem{c, r, static_cast<float>(values[o])});
        }
    }
    // schedule every list; PEs in parallel
    std::vector<std::vector<Elem>> sched(bucket.size());
    const int groups = n / (2 * kCuperPes) + 1;
    const int workers = static_cast<int>(std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency())));
    for_workers(workers, [&](int w) {
        std::vector<int> window(static_cast<size_t>(groups), -kCuperWindow);
        std::vector<int> touched;
        for (int pe = w; pe < kCuperPes; pe += workers)
            for (int b = 0; b < batches; ++b) {
                const size_t k = static_cast<size_t>(b) * kCuperPes + static_cast<size_t>(pe);
                schedule(bucket[k], b * kCuperSliceWidth, window, touched, sched[k]);
                std::vector<Elem>().swap(bucket[k]);
            }
    });
    // every batch padded to its longest PE list
    for (int b = 0; b < batches; ++b) {
        size_t longest = 0;
        for (int pe = 0; pe < kCuperPes; ++pe)
            longest = std::max(longest, sched[static_cast<size_t>(b) * kCuperPes + static_cast<size_t>(pe)].size());
        p.ptr[static_cast<size_t>(b) + 1] = p.ptr[static_cast<size_t>(b)] + static_cast<int>(longest);
    }
    p.matrix_len = p.ptr.back();
    p.channel_size = (8 * p.matrix_len + 511) / 512 * 512;
    p.data.assign(static_cast<size_t>(kCuperChannels) * static_cast<size_t>(p.channel_size), 0);
    const Elem empty;
    for_workers(workers, [&](int w) {
        for (int pe = w; pe < kCuperPes; pe += workers) {
            const int c = pe / kCuperLanes, lane = pe % kCuperLanes;
            uint64_t* ch = p.data.data() + static_cast<size_t>(c) * static_cast<size_t>(p.channel_size);
            for (int b = 0; b < batches; ++b) {
                const std::vector<Elem>& s = sched[static_cast<size_t>(b) * kCuperPes + static_cast<size_t>(pe)];
                for (int i = p.ptr[static_cast<size_t>(b)]; i < p.ptr[static_cast<size_t>(b) + 1]; ++i) {
                    const size_t k = static_cast<size_t>(i - p.ptr[static_cast<size_t>(b)]);
                    ch[static_cast<size_t>(i) * 8 + static_cast<size_t>(lane)] = pack_word(k < s.size() ? s[k] : empty);
                }
            }
        }
    });
    return p;
}

HostCsr cuper_unpack(const CuperPacked& p, int n) {
    if (n < 0 || n > kCuperMaxRows) throw std::runtime_error("n out of range for the Cuper format");
    if (p.batch_count < 0 || p.ptr.size() != static_cast<size_t>(p.batch_count) + 1 || p.ptr[0] != 0)
        throw std::runtime_error("malformed sp_element_list_ptr");
    for (int b = 0; b < p.batch_count; ++b)
        if (p.ptr[static_cast<size_t>(b) + 1] < p.ptr[static_cast<size_t>(b)])
            throw std::runtime_error("sp_element_list_ptr must be nondecreasing");
    if (p.matrix_len != p.ptr.back() || p.channel_size < 8 * static_cast<long long>(p.matrix_len) ||
        p.data.size() != static_cast<size_t>(kCuperChannels) * static_cast<size_t>(p.channel_size))
        throw std::runtime_error("matrix_data does not match matrix_len / channel_size");
    struct Coo {
        int row, col;
        double v;
    };
    std::vector<Coo> coo;
    for (int c = 0; c < kCuperChannels; ++c) {
        const uint64_t* ch = p.data.data() + static_cast<size_t>(c) * static_cast<size_t>(p.channel_size);
        for (int lane = 0; lane < kCuperLanes; ++lane)
            for (int b = 0; b < p.batch_count; ++b)
                for (int i = p.ptr[static_cast<size_t>(b)]; i < p.ptr[static_cast<size_t>(b) + 1]; ++i) {
                    const uint64_t w = ch[static_cast<size_t>(i) * 8 + static_cast<size_t>(lane)];
                    const int code = static_cast<int>((w >> 32) & 0x3FFFFull);
                    if (code & (1 << 17)) continue;   // empty slot
                    // restore_global_row (kernels/cuper_pcg_control_kernel.cpp:140-147)
                    const long long packet = static_cast<long long>(code >> 1) * kCuperPes + c / 2 + (c & 1) * 8 +
                                             lane * 16;
                    const long long row = packet * 2 + (code & 1);
                    const long long col = static_cast<long long>(b) * kCuperSliceWidth + static_cast<long long>(w >> 50);
                    if (row >= n || col >= n)
                        throw std::runtime_error("packed entry outside the " + std::to_string(n) + " x " +
                                                 std::to_string(n) + " matrix");
                    float f;
                    const uint32_t bits = static_cast<uint32_t>(w & 0xFFFFFFFFull);
                    std::memcpy(&f, &bits, 4);
                    coo.push_back(Coo{static_cast<int>(row), static_cast<int>(col), static_cast<double>(f)});
                }
    }
    std::stable_sort(coo.begin(), coo.end(),
                     [](const Coo& a, const Coo& b) { return a.row != b.row ? a.row < b.row : a.col < b.col; });
    HostCsr a;
    a.n = n;
    a.row_ptr.assign(static_cast<size_t>(n) + 1, 0);
    a.col_idx.resize(coo.size());
    a.values.resize(coo.size());
    for (size_t k = 0; k < coo.size(); ++k) {
        ++a.row_ptr[static_cast<size_t>(coo[k].row) + 1];
        a.col_idx[k] = coo[k].col;
        a.values[k] = coo[k].v;
    }
    for (int r = 0; r < n; ++r) a.row_ptr[static_cast<size_t>(r) + 1] += a.row_ptr[static_cast<size_t>(r)];
    a.b.assign(static_cast<size_t>(n), 0.0);
    a.x0.assign(static_cast<size_t>(n), 0.0);
    return a;
}

}  // namespace xp

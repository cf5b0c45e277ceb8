// mtx_ingest.cpp -- MatrixMarket ingestion (parallel parse) and the binary
// CSR cache.  Semantics are listed in mtx_ingest.hpp.
#include "mtx_ingest.hpp"

#include <algorithm>
#include <cctype>
#include <charconv>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <filesystem>
#include <stdexcept>
#include <thread>

#include <unistd.h>

#include "csr_dataset.hpp"

namespace xp {

namespace fs = std::filesystem;

namespace {

std::string read_whole_file(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (f == nullptr) throw std::runtime_error("failed to open " + path);
    std::string s;
    char buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, got);
    const bool bad = std::ferror(f) != 0;
    std::fclose(f);
    if (bad) throw std::runtime_error("failed to read " + path);
    return s;
}

// Python's str.strip()/split() whitespace, minus the newline that ends a line.
inline bool is_ws(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

// Next line of [p, end) with surrounding whitespace removed; p moves past it.
inline bool next_line(const char*& p, const char* end, const char*& lb, const char*& le) {
    if (p >= end) return false;
    const char* nl = static_cast<const char*>(std::memchr(p, '\n', static_cast<size_t>(end - p)));
    const char* e = nl != nullptr ? nl : end;
    lb = p;
    le = e;
    p = nl != nullptr ? nl + 1 : end;
    while (lb < le && is_ws(*lb)) ++lb;
    while (le > lb && is_ws(le[-1])) --le;
    return true;
}

inline bool next_token(const char*& p, const char* e, const char*& tb, const char*& te) {
    while (p < e && is_ws(*p)) ++p;
    if (p >= e) return false;
    tb = p;
    while (p < e && !is_ws(*p)) ++p;
    te = p;
    return true;
}

// Python int(): optional sign then decimal digits.
bool parse_int(const char* b, const char* e, long long& out) {
    bool neg = false;
    if (b < e && (*b == '+' || *b == '-')) neg = *b++ == '-';
    if (b == e) return false;
    long long v = 0;
    for (; b < e; ++b) {
        if (*b < '0' || *b > '9' || v > (LLONG_MAX - 9) / 10) return false;
        v = v * 10 + (*b - '0');
    }
    out = neg ? -v : v;
    return true;
}

// Python float() on one token: decimal or exponent form, inf/nan, optional
// sign; correctly rounded.  Hexadecimal forms are rejected, as Python does.
bool parse_double(const char* b, const char* e, double& out) {
    const char* s = b;
    if (s < e && *s == '+') {
        ++s;
        if (s < e && (*s == '-' || *s == '+')) return false;
    }
    const auto r = std::from_chars(s, e, out, std::chars_format::general);
    if (r.ec == std::errc::result_out_of_range && r.ptr == e) {
        const std::string t(b, e);     // overflow to +-inf, underflow to 0 or a subnormal
        char* endp = nullptr;
        out = std::strtod(t.c_str(), &endp);
        return endp == t.c_str() + t.size();
    }
    return r.ec == std::errc() && r.ptr == e;
}

std::string lower(std::string s) {
    for (char& c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    return s;
}

struct MmFile {
    std::string text;
    std::vector<std::string> header;   // the banner tokens
    std::string size_line;             // the first data line
    const char* body = nullptr;        // first byte after the size line
    const char* end = nullptr;
};

// read_mm_header + next_data_line (download_suitesparse_data.py:197-211).
MmFile open_mm(const std::string& path) {
    MmFile f;
    f.text = read_whole_file(path);
    const char* p = f.text.data();
    f.end = p + f.text.size();
    const char *lb, *le, *tb, *te;
    if (next_line(p, f.end, lb, le))
        while (next_token(lb, le, tb, te)) f.header.emplace_back(tb, te);
    if (f.header.size() < 5 || f.header[0] != "%%MatrixMarket" || f.header[1] != "matrix")
        throw std::runtime_error("unsupported MatrixMarket header in " + path);
    while (true) {
        if (!next_line(p, f.end, lb, le)) throw std::runtime_error("unexpected end of MatrixMarket file");
        if (lb < le && *lb != '%') break;
    }
    f.size_line.assign(lb, le);
    f.body = p;
    return f;
}

// The size line's leading `count` integers.
std::vector<long long> size_line(const MmFile& f, int count, const std::string& path) {
    std::vector<long long> dims;
    const char* p = f.size_line.data();
    const char* e = p + f.size_line.size();
    const char *tb, *te;
    for (int k = 0; k < count; ++k) {
        long long v;
        if (!next_token(p, e, tb, te) || !parse_int(tb, te, v))
            throw std::runtime_error("invalid MatrixMarket size line in " + path);
        dims.push_back(v);
    }
    return dims;
}

struct Entry {
    int r, c;
    double v;
};

struct Chunk {
    std::vector<Entry> entries;
    long long lines = 0;
    std::string error;     // first malformed line of this chunk
};

// One chunk of entry lines, kept in file order: (i, j) when inside the
// subset, then its mirror (j, i) for an off-diagonal symmetric entry.
void parse_chunk(const char* p, const char* end, bool pattern, bool symmetric, long long out_rows,
                 long long out_cols, Chunk& out) {
    const char *lb, *le, *tb, *te;
    while (next_line(p, end, lb, le)) {
        if (lb == le || *lb == '%') continue;
        ++out.lines;
        const char* q = lb;
        long long i = 0, j = 0;
        double v = 1.0;
        bool ok = next_token(q, le, tb, te) && parse_int(tb, te, i);
        ok = ok && next_token(q, le, tb, te) && parse_int(tb, te, j);
        if (ok && !pattern) ok = next_token(q, le, tb, te) && parse_double(tb, te, v);
        if (!ok) {
            out.error = "malformed MatrixMarket entry '" + std::string(lb, le) + "'";
            return;
        }
        if (i < 1 || j < 1) {
            out.error = "MatrixMarket index below 1 in entry '" + std::string(lb, le) + "'";
            return;
        }
        --i;
        --j;
        if (v == 0.0) continue;
        if (i < out_rows && j < out_cols) out.entries.push_back({static_cast<int>(i), static_cast<int>(j), v});
        if (symmetric && i != j && j < out_rows && i < out_cols)
            out.entries.push_back({static_cast<int>(j), static_cast<int>(i), v});
    }
}

int worker_count(size_t bytes) {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t by_size = bytes / (1u << 20) + 1;
    return static_cast<int>(std::min<size_t>({static_cast<size_t>(hw), by_size, 16}));
}

// f(0) .. f(workers - 1), one thread each.  A worker's exception is
// rethrown after every thread has joined; a thread that cannot be started
// runs its share inline.
template <typename F>
void parallel_for(int workers, F&& f) {
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

MtxMatrix read_mtx_matrix(const std::string& path, long long subset_n) {
    MmFile f = open_mm(path);
    const std::string storage = lower(f.header[2]), field = lower(f.header[3]), symmetry = lower(f.header[4]);
    if (storage != "coordinate") throw std::runtime_error("expected coordinate MatrixMarket matrix: " + path);
    if (field != "real" && field != "integer" && field != "pattern")
        throw std::runtime_error("unsupported MatrixMarket field '" + field + "' in " + path);
    if (symmetry != "general" && symmetry != "symmetric")
        throw std::runtime_error("unsupported MatrixMarket symmetry '" + symmetry + "' in " + path);

    const std::vector<long long> dims = size_line(f, 3, path);
    MtxMatrix M;
    M.nrows = dims[0];
    M.ncols = dims[1];
    M.nnz_reported = dims[2];
    M.field = field == "real" ? 0 : field == "integer" ? 1 : 2;
    M.symmetry = symmetry == "symmetric" ? 1 : 0;
    const long long out_rows = subset_n > 0 ? std::min(M.nrows, subset_n) : M.nrows;
    const long long out_cols = subset_n > 0 ? std::min(M.ncols, subset_n) : M.ncols;
    if (out_rows < 0 || out_cols < 0 || out_rows > INT_MAX || out_cols > INT_MAX)
        throw std::runtime_error("MatrixMarket shape out of range in " + path);
    M.n = static_cast<int>(out_rows);
    M.m = static_cast<int>(out_cols);

    // Parse: chunks split at line ends, one thread each.
    const size_t body_bytes = static_cast<size_t>(f.end - f.body);
    const int T = worker_count(body_bytes);
    std::vector<const char*> cut(static_cast<size_t>(T) + 1);
    cut[0] = f.body;
    cut[static_cast<size_t>(T)] = f.end;
    for (int k = 1; k < T; ++k) {
        const char* c = f.body + body_bytes * static_cast<size_t>(k) / static_cast<size_t>(T);
        c = std::max(c, cut[static_cast<size_t>(k) - 1]);
        const char* nl = static_cast<const char*>(std::memchr(c, '\n', static_cast<size_t>(f.end - c)));
        cut[static_cast<size_t>(k)] = nl != nullptr ? nl + 1 : f.end;
    }
    std::vector<Chunk> chunks(static_cast<size_t>(T));
    parallel_for(T, [&](int w) {
        parse_chunk(cut[static_cast<size_t>(w)], cut[static_cast<size_t>(w) + 1], M.field == 2, M.symmetry == 1,
                    out_rows, out_cols, chunks[static_cast<size_t>(w)]);
    });
    size_t total = 0;
    for (const Chunk& c : chunks) {
        if (!c.error.empty()) throw std::runtime_error(c.error + " in " + path);
        M.entries += c.lines;
        total += c.entries.size();
    }
    if (total > static_cast<size_t>(INT_MAX)) throw std::runtime_error("too many entries for 32-bit CSR in " + path);
    f.text.clear();
    f.text.shrink_to_fit();

    // Stable counting sort by row: inside a row the entries stay in file order.
    const size_t n = static_cast<size_t>(M.n);
    std::vector<int> start(n + 1, 0);
    for (const Chunk& c : chunks)
        for (const Entry& e : c.entries) ++start[static_cast<size_t>(e.r) + 1];
    for (size_t r = 0; r < n; ++r) start[r + 1] += start[r];
    std::vector<int> cols(total);
    std::vector<double> vals(total);
    {
        std::vector<int> pos(start.begin(), start.end() - 1);
        for (Chunk& c : chunks) {
            for (const Entry& e : c.entries) {
                const int k = pos[static_cast<size_t>(e.r)]++;
                cols[static_cast<size_t>(k)] = e.c;
                vals[static_cast<size_t>(k)] = e.v;
            }
            std::vector<Entry>().swap(c.entries);
        }
    }

    // Per row: stable sort by column, sum duplicates in file order, then drop
    // the sums that are exactly zero (parse_coordinate_matrix :266-277).
    std::vector<int> kept(n, 0);
    const int W = std::min<int>(T, static_cast<int>(std::max<size_t>(1, n / 4096)));
    parallel_for(W, [&](int w) {
        std::vector<std::pair<int, double>> seg;
        const size_t r0 = n * static_cast<size_t>(w) / static_cast<size_t>(W);
        const size_t r1 = n * (static_cast<size_t>(w) + 1) / static_cast<size_t>(W);
        for (size_t r = r0; r < r1; ++r) {
            const size_t a = static_cast<size_t>(start[r]), b = static_cast<size_t>(start[r + 1]);
            seg.clear();
            for (size_t k = a; k < b; ++k) seg.emplace_back(cols[k], vals[k]);
            std::stable_sort(seg.begin(), seg.end(),
                             [](const std::pair<int, double>& x, const std::pair<int, double>& y) {
                                 return x.first < y.first;
                             });
            size_t w2 = 0;
            for (size_t k = 0; k < seg.size(); ++k) {
                if (w2 > 0 && seg[w2 - 1].first == seg[k].first)
                    seg[w2 - 1].second += seg[k].second;
                else
                    seg[w2++] = seg[k];
            }
            size_t out = a;
            for (size_t k = 0; k < w2; ++k) {
                if (seg[k].second == 0.0) continue;
                cols[out] = seg[k].first;
                vals[out] = seg[k].second;
                ++out;
            }
            kept[r] = static_cast<int>(out - a);
        }
    });

    M.row_ptr.assign(n + 1, 0);
    for (size_t r = 0; r < n; ++r) M.row_ptr[r + 1] = M.row_ptr[r] + kept[r];
    const size_t nnz = static_cast<size_t>(M.row_ptr[n]);
    M.col_idx.resize(nnz);
    M.values.resize(nnz);
    for (size_t r = 0; r < n; ++r) {
        const size_t a = static_cast<size_t>(start[r]), d = static_cast<size_t>(M.row_ptr[r]);
        std::copy_n(cols.begin() + static_cast<long>(a), kept[r], M.col_idx.begin() + static_cast<long>(d));
        std::copy_n(vals.begin() + static_cast<long>(a), kept[r], M.values.begin() + static_cast<long>(d));
        for (int k = 0; k < kept[r]; ++k) {
            if (M.col_idx[d + static_cast<size_t>(k)] != static_cast<int>(r)) continue;
            const double av = std::fabs(M.values[d + static_cast<size_t>(k)]);
            M.diag_min_abs = M.has_diag ? std::min(M.diag_min_abs, av) : av;
            M.has_diag = true;
        }
    }
    return M;
}

std::vector<double> read_mtx_vector(const std::string& path, int expected_n, long long subset_n) {
    const MmFile f = open_mm(path);
    const std::string storage = lower(f.header[2]), field = lower(f.header[3]);
    if (field != "real" && field != "integer" && field != "pattern")
        throw std::runtime_error("unsupported vector field '" + field + "' in " + path);
    const long long limit = subset_n > 0 ? std::min<long long>(expected_n, subset_n) : expected_n;
    std::vector<double> vec(static_cast<size_t>(std::max<long long>(limit, 0)), 0.0);
    const char *lb, *le, *tb, *te;
    const char* p = f.body;

    if (storage == "array") {
        const std::vector<long long> dims = size_line(f, 2, path);
        if (dims[0] < limit || dims[1] < 1) throw std::runtime_error("vector dimensions do not match dataset: " + path);
        long long idx = 0;
        while (idx < limit && next_line(p, f.end, lb, le)) {
            if (lb == le || *lb == '%') continue;
            while (idx < limit && next_token(lb, le, tb, te)) {
                double v;
                if (!parse_double(tb, te, v))
                    throw std::runtime_error("malformed vector entry '" + std::string(tb, te) + "' in " + path);
                vec[static_cast<size_t>(idx++)] = v;
            }
        }
        if (idx < limit) throw std::runtime_error("not enough vector entries in " + path);
        return vec;
    }
    if (storage == "coordinate") {
        const std::vector<long long> dims = size_line(f, 3, path);
        if (dims[0] < limit || dims[1] < 1) throw std::runtime_error("vector dimensions do not match dataset: " + path);
        while (next_line(p, f.end, lb, le)) {
            if (lb == le || *lb == '%') continue;
            const char* q = lb;
            long long i = 0, j = 0;
            if (!(next_token(q, le, tb, te) && parse_int(tb, te, i) && next_token(q, le, tb, te) &&
                  parse_int(tb, te, j)))
                throw std::runtime_error("malformed vector entry '" + std::string(lb, le) + "' in " + path);
            if (i < 1 || j < 1)
                throw std::runtime_error("MatrixMarket index below 1 in entry '" + std::string(lb, le) + "' in " +
                                         path);
            if (i - 1 >= limit || j - 1 != 0) continue;
            double v = 1.0;
            if (field != "pattern" && !(next_token(q, le, tb, te) && parse_double(tb, te, v)))
                throw std::runtime_error("malformed vector entry '" + std::string(lb, le) + "' in " + path);
            vec[static_cast<size_t>(i - 1)] = v;
        }
        return vec;
    }
    throw std::runtime_error("unsupported vector MatrixMarket storage '" + storage + "' in " + path);
}

std::vector<double> reference_solution(int n) {
#pragma clang fp contract(off)
    // x_i = 0.75 + 0.18 sin(0.013 i) + 0.12 cos(0.017 i) + 9e-7 i
    //       + 0.28 exp(-((i - c) / w)^2),  c = 0.63 max(1, n-1), w = max(8, 0.07 n)
    // (download_suitesparse_data.py:339-349), summed left to right with the
    // C library's sin, cos, exp and pow, as the converter evaluates it.
    std::vector<double> x(static_cast<size_t>(std::max(n, 0)));
    const double center = static_cast<double>(std::max(1, n - 1)) * 0.63;
    const double width = std::max(8.0, static_cast<double>(n) * 0.07);
    volatile double two = 2.0;     // a run-time pow(t, 2), not a folded t * t
    const double e2 = two;
    for (int i = 0; i < n; ++i) {
        const double di = static_cast<double>(i);
        const double t = (di - center) / width;
        double v = 0.75 + 0.18 * std::sin(0.013 * di);
        v = v + 0.12 * std::cos(0.017 * di);
        v = v + 0.0000009 * di;
        v = v + 0.28 * std::exp(-std::pow(t, e2));
        x[static_cast<size_t>(i)] = v;
    }
    return x;
}

void csr_spmv_host_sequential(const HostCsr& a, const std::vector<double>& x, std::vector<double>& y) {
#pragma clang fp contract(off)
    y.assign(static_cast<size_t>(a.n), 0.0);
    const int W = std::min(worker_count(a.values.size() * 12), std::max(1, a.n / 4096));
    parallel_for(W, [&](int w) {
        const int r0 = static_cast<int>(static_cast<long long>(a.n) * w / W);
        const int r1 = static_cast<int>(static_cast<long long>(a.n) * (w + 1) / W);
        for (int r = r0; r < r1; ++r) {
            double acc = 0.0;
            for (int o = a.row_ptr[static_cast<size_t>(r)]; o < a.row_ptr[static_cast<size_t>(r) + 1]; ++o)
                acc = acc + a.values[static_cast<size_t>(o)] * x[static_cast<size_t>(a.col_idx[static_cast<size_t>(o)])];
            y[static_cast<size_t>(r)] = acc;
        }
    });
}

HostCsr mtx_dataset(const std::string& mtx_path, const std::string& b_path, long long subset_n, MtxMatrix* info) {
    MtxMatrix M = read_mtx_matrix(mtx_path, subset_n);
    if (M.n != M.m)
        throw std::runtime_error(fs::path(mtx_path).stem().string() + " is not square after conversion: " +
                                 std::to_string(M.n) + "x" + std::to_string(M.m));
    HostCsr a;
    a.n = M.n;
    a.row_ptr = std::move(M.row_ptr);
    a.col_idx = std::move(M.col_idx);
    a.values = std::move(M.values);
    if (!b_path.empty())
        a.b = read_mtx_vector(b_path, a.n, subset_n);
    else
        csr_spmv_host_sequential(a, reference_solution(a.n), a.b);
    a.x0.assign(static_cast<size_t>(a.n), 0.0);
    if (info != nullptr) {
        *info = std::move(M);
        info->row_ptr.clear();
        info->col_idx.clear();
        info->values.clear();
    }
    return a;
}

// ---- binary cache -----------------------------------------------------------
namespace {

constexpr char kMagic[8] = {'X', 'P', 'C', 'S', 'R', 'B', 'I', 'N'};
constexpr uint32_t kVersion = 1;
const char* const kTextFiles[5] = {"row_ptr.txt", "col_idx.txt", "values.txt", "b.txt", "x0.txt"};

struct BinHeader {
    char magic[8];
    uint32_t version;
    uint32_t header_bytes;
    int64_t n;
    int64_t nnz;
    uint64_t checksum;
    uint64_t src_size[5];
    int64_t src_mtime_ns[5];
    uint64_t reserved;
};
static_assert(sizeof(BinHeader) == 128, "cache header is 128 bytes");

template <typename T>
uint64_t chain(uint64_t h, const std::vector<T>& v) {
    return payload_checksum(v.data(), v.size() * sizeof(T), h);
}

uint64_t checksum_of(const HostCsr& a) {
    uint64_t h = 0x243F6A8885A308D3ull;
    h = chain(h, a.row_ptr);
    h = chain(h, a.col_idx);
    h = chain(h, a.values);
    h = chain(h, a.b);
    h = chain(h, a.x0);
    return h;
}

template <typename T>
void write_vec(FILE* f, const std::vector<T>& v, const std::string& path) {
    if (!v.empty() && std::fwrite(v.data(), sizeof(T), v.size(), f) != v.size())
        throw std::runtime_error("failed to write " + path);
}

template <typename T>
void read_vec(FILE* f, std::vector<T>& v, size_t count, const std::string& path) {
    v.resize(count);
    if (count != 0 && std::fread(v.data(), sizeof(T), count, f) != count)
        throw std::runtime_error("dataset cache truncated: " + path);
}

}  // namespace

uint64_t payload_checksum(const void* data, size_t bytes, uint64_t h) {
    // Four independent multiply-xorshift lanes over 8-byte words, folded
    // together with the length; a bit flip anywhere changes the result.
    constexpr uint64_t K = 0x9E3779B97F4A7C15ull;
    uint64_t l[4] = {h ^ 0x13198A2E03707344ull, h + K, h ^ 0xA4093822299F31D0ull, ~h};
    const unsigned char* p = static_cast<const unsigned char*>(data);
    const size_t words = bytes / 8;
    size_t i = 0;
    for (; i + 4 <= words; i += 4) {
        for (int k = 0; k < 4; ++k) {
            uint64_t w;
            std::memcpy(&w, p + 8 * (i + static_cast<size_t>(k)), 8);
            l[k] = (l[k] ^ w) * K;
            l[k] ^= l[k] >> 29;
        }
    }
    for (; i < words; ++i) {
        uint64_t w;
        std::memcpy(&w, p + 8 * i, 8);
        l[0] = (l[0] ^ w) * K;
        l[0] ^= l[0] >> 29;
    }
    uint64_t tail = 0;
    std::memcpy(&tail, p + 8 * words, bytes - 8 * words);
    uint64_t r = (l[0] ^ tail ^ static_cast<uint64_t>(bytes)) * K;
    for (int k = 1; k < 4; ++k) {
        r = (r ^ l[k]) * K;
        r ^= r >> 31;
    }
    return r;
}

bool SourceStamp::operator==(const SourceStamp& o) const {
    for (int k = 0; k < 5; ++k)
        if (size[k] != o.size[k] || mtime_ns[k] != o.mtime_ns[k]) return false;
    return true;
}

SourceStamp stamp_dataset_dir(const std::string& dir) {
    SourceStamp s;
    for (int k = 0; k < 5; ++k) {
        const fs::path p = fs::path(dir) / kTextFiles[k];
        s.size[k] = static_cast<uint64_t>(fs::file_size(p));
        s.mtime_ns[k] = static_cast<int64_t>(
            std::chrono::duration_cast<std::chrono::nanoseconds>(fs::last_write_time(p).time_since_epoch()).count());
    }
    return s;
}

void save_binary(const HostCsr& a, const std::string& path, const SourceStamp& stamp) {
    // Validated like a text dataset before it is written.
    CsrDataset::from_arrays(a.row_ptr, a.col_idx, a.values, a.b, a.x0);
    BinHeader h{};
    std::memcpy(h.magic, kMagic, 8);
    h.version = kVersion;
    h.header_bytes = sizeof(BinHeader);
    h.n = a.n;
    h.nnz = static_cast<int64_t>(a.col_idx.size());
    h.checksum = checksum_of(a);
    for (int k = 0; k < 5; ++k) {
        h.src_size[k] = stamp.size[k];
        h.src_mtime_ns[k] = stamp.mtime_ns[k];
    }
    // Written under a temporary name and renamed, so a reader never sees a
    // partial file.
    const std::string tmp = path + ".tmp." + std::to_string(static_cast<long long>(getpid()));
    FILE* f = std::fopen(tmp.c_str(), "wb");
    if (f == nullptr) throw std::runtime_error("failed to open " + tmp);
    try {
        if (std::fwrite(&h, sizeof h, 1, f) != 1) throw std::runtime_error("failed to write " + tmp);
        write_vec(f, a.row_ptr, tmp);
        write_vec(f, a.col_idx, tmp);
        write_vec(f, a.values, tmp);
        write_vec(f, a.b, tmp);
        write_vec(f, a.x0, tmp);
        if (std::fclose(f) != 0) {
            f = nullptr;
            throw std::runtime_error("failed to write " + tmp);
        }
        f = nullptr;
        fs::rename(tmp, path);
    } catch (...) {
        if (f != nullptr) std::fclose(f);
        std::error_code ec;
        fs::remove(tmp, ec);
        throw;
    }
}

HostCsr load_binary(const std::string& path, SourceStamp* stamp) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (f == nullptr) throw std::runtime_error("failed to open " + path);
    HostCsr a;
    try {
        BinHeader h{};
        if (std::fread(&h, sizeof h, 1, f) != 1 || std::memcmp(h.magic, kMagic, 8) != 0 ||
            h.header_bytes != sizeof(BinHeader))
            throw std::runtime_error("not a dataset cache file: " + path);
        if (h.version != kVersion) throw std::runtime_error("unsupported dataset cache version in " + path);
        if (h.n < 0 || h.n > INT_MAX - 1 || h.nnz < 0 || h.nnz > INT_MAX)
            throw std::runtime_error("dataset cache header out of range: " + path);
        const uint64_t want = sizeof(BinHeader) + static_cast<uint64_t>(h.n + 1) * 4 + static_cast<uint64_t>(h.nnz) * 12 +
                              static_cast<uint64_t>(h.n) * 16;
        if (static_cast<uint64_t>(fs::file_size(path)) != want)
            throw std::runtime_error("dataset cache truncated: " + path);
        a.n = static_cast<int>(h.n);
        read_vec(f, a.row_ptr, static_cast<size_t>(h.n) + 1, path);
        read_vec(f, a.col_idx, static_cast<size_t>(h.nnz), path);
        read_vec(f, a.values, static_cast<size_t>(h.nnz), path);
        read_vec(f, a.b, static_cast<size_t>(h.n), path);
        read_vec(f, a.x0, static_cast<size_t>(h.n), path);
        if (checksum_of(a) != h.checksum) throw std::runtime_error("dataset cache checksum mismatch: " + path);
        if (stamp != nullptr)
            for (int k = 0; k < 5; ++k) {
                stamp->size[k] = h.src_size[k];
                stamp->mtime_ns[k] = h.src_mtime_ns[k];
            }
    } catch (...) {
        std::fclose(f);
        throw;
    }
    std::fclose(f);
    // The text loader's structural checks (row_ptr monotone, columns in range).
    CsrDataset::from_arrays(a.row_ptr, a.col_idx, a.values, a.b, a.x0);
    return a;
}

HostCsr load_cached(const std::string& dir, const std::string& cache_path, bool* hit) {
    if (hit != nullptr) *hit = false;
    const std::string cache = cache_path.empty() ? (fs::path(dir) / "csr_cache.xpbin").string() : cache_path;
    SourceStamp now;
    bool have_stamp = true;
    try {
        now = stamp_dataset_dir(dir);
    } catch (const std::exception&) {
        have_stamp = false;      // a missing file: the text loader reports it below
    }
    if (have_stamp && fs::exists(cache)) {
        try {
            SourceStamp was;
            HostCsr a = load_binary(cache, &was);
            if (was == now) {
                if (hit != nullptr) *hit = true;
                return a;
            }
        } catch (const std::exception&) {
            // unreadable or stale cache: rebuilt from the text files
        }
    }
    const CsrDataset d = CsrDataset::load(dir);
    HostCsr a;
    a.n = d.n();
    a.row_ptr = d.row_ptr();
    a.col_idx = d.col_idx();
    a.values = d.values();
    a.b = d.b();
    a.x0 = d.x0();
    try {
        save_binary(a, cache, now);
    } catch (const std::exception&) {
        // read-only location: the dataset is still returned
    }
    return a;
}

}  // namespace xp

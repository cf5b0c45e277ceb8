// device_build.hip -- matrix set-up kernels: CSR validation and row facts,
// the parallel strength aggregation (blocking rule 3) and the SELL-C-sigma
// build, all from a CSR resident in device memory.  See device_build.hpp.
//
// Everything here is integer / byte work plus a few FP64 sums whose order
// is fixed (CSR order within a row), so results are deterministic and equal
// to the host code's: products and sums are kept apart (no fused
// multiply-add) where the host rounds them separately.  Reductions across
// rows use integer atomics only (max of packed keys, counts, min of row
// indices), whose results do not depend on the order of arrival.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "device_build.hpp"

namespace xp {

namespace {

constexpr int kRowThreads = 256;
constexpr int kScanThreads = 1024;
constexpr int kScanPer = 4;
constexpr int kScanTile = kScanThreads * kScanPer;
constexpr int kSortThreads = 1024;

int blocks_of(long long n, int per) { return static_cast<int>((n + per - 1) / per); }

// Work buffers of the set-up, kept between builds and grown on demand: a
// build then makes no hipMalloc / hipFree calls for its temporaries (about
// 100 bytes per row), only for the arrays the matrix keeps.  Builds are
// serialised by set_up_mutex().
enum Work : int {
    kWErr, kWOwner, kWSize, kWMembers, kWM1, kWCounter, kWTaken, kWBid, kWElig, kWFlagUnit, kWUnitSrc, kWFlagPiece,
    kWPieceSrc, kWPos, kWScan, kWTotals, kWPieces, kWPsz, kWPdst, kWProws, kWPsize, kWKey, kWOrder, kWWidth,
    kWFlags, kWSmin, kWSmax, kWOffcnt, kWDiagPos, kWDis, kWTotal, kWCount
};

DevBuf<unsigned char>& work_slot(int id) {
    static DevBuf<unsigned char> slots[kWCount];
    return slots[id];
}

template <typename T>
T* take(int id, size_t count) {
    DevBuf<unsigned char>& b = work_slot(id);
    b.ensure(std::max<size_t>(count, 1) * sizeof(T));
    return reinterpret_cast<T*>(b.p);
}

// ---- exclusive scan of int counts ----------------------------------------------
__device__ __forceinline__ int wave_incl_scan(int v) {
    const int lane = static_cast<int>(threadIdx.x) & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

// Exclusive scan over a 1024-thread workgroup; *total gets the sum.
__device__ int block_excl_scan(int v, int* total) {
    __shared__ int ws[kScanThreads / 64];
    const int tid = static_cast<int>(threadIdx.x);
    const int wave = tid >> 6, lane = tid & 63;
    const int inc = wave_incl_scan(v);
    if (lane == 63) ws[wave] = inc;
    __syncthreads();
    if (wave == 0) {
        int x = lane < kScanThreads / 64 ? ws[lane] : 0;
        x = wave_incl_scan(x);
        if (lane < kScanThreads / 64) ws[lane] = x;
    }
    __syncthreads();
    const int off = wave > 0 ? ws[wave - 1] : 0;
    *total = ws[kScanThreads / 64 - 1];
    __syncthreads();  // ws is reused by the next call
    return off + inc - v;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_tile(int n, const int* __restrict__ in, int* __restrict__ out,
                                                            int* __restrict__ tile_sums) {
    const long long base = static_cast<long long>(blockIdx.x) * kScanTile + static_cast<long long>(threadIdx.x) * kScanPer;
    int v[kScanPer];
    int sum = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        v[k] = base + k < n ? in[base + k] : 0;
        sum += v[k];
    }
    int total;
    int run = block_excl_scan(sum, &total);
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_tiles(int ntiles, int* __restrict__ tile_sums,
                                                             int* __restrict__ total_out) {
    int carry = 0;
    for (int i0 = 0; i0 < ntiles; i0 += kScanThreads) {
        const int i = i0 + static_cast<int>(threadIdx.x);
        const int v = i < ntiles ? tile_sums[i] : 0;
        int tot;
        const int ex = block_excl_scan(v, &tot);
        if (i < ntiles) tile_sums[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0 && total_out != nullptr) *total_out = carry;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_add(int n, int* __restrict__ out,
                                                           const int* __restrict__ tile_sums) {
    const long long i0 = static_cast<long long>(blockIdx.x) * kScanTile;
    const int add = tile_sums[blockIdx.x];
    for (int k = static_cast<int>(threadIdx.x); k < kScanTile; k += kScanThreads)
        if (i0 + k < n) out[i0 + k] += add;
}

// out[i] = sum in[0..i), *total = sum in[0..n) (device pointer, may be null).
void scan_exclusive(const int* in, int* out, int n, int* total, hipStream_t s) {
    const int ntiles = std::max(1, blocks_of(n, kScanTile));
    int* tmp = take<int>(kWScan, static_cast<size_t>(ntiles));
    hipLaunchKernelGGL(k_scan_tile, dim3(ntiles), dim3(kScanThreads), 0, s, n, in, out, tmp);
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kScanThreads), 0, s, ntiles, tmp, total);
    hipLaunchKernelGGL(k_scan_add, dim3(ntiles), dim3(kScanThreads), 0, s, n, out, tmp);
}

template <typename T>
__global__ void k_fill_n(long long n, T v, T* __restrict__ dst) {
    const long long i = static_cast<long long>(blockIdx.x) * kRowThreads + threadIdx.x;
    if (i < n) dst[i] = v;
}

template <typename T>
void fill(T* p, long long n, T v, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_fill_n<T>, dim3(blocks_of(n, kRowThreads)), dim3(kRowThreads), 0, s, n, v, p);
}

// Sum of one value over a 256-thread workgroup, then one atomic per group.
__device__ __forceinline__ void group_count(int v, int* counter) {
    __shared__ int part[kRowThreads / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int wave = static_cast<int>(threadIdx.x) >> 6;
    if ((threadIdx.x & 63) == 0) part[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < kRowThreads / 64; ++w) t += part[w];
        if (t != 0) atomicAdd(counter, t);
    }
}

// ---- row facts ----------------------------------------------------------------------
enum : int { kErrRowPtr = 2, kErrCol = 4 };

__global__ void k_row_info(int n, long long nnz, const int* __restrict__ rp, const int* __restrict__ ci,
                           const double* __restrict__ va, int* __restrict__ offcnt, int* __restrict__ diag_pos,
                           double* __restrict__ dis, int* __restrict__ err) {
    const int r = static_cast<int>(blockIdx.x) * kRowThreads + static_cast<int>(threadIdx.x);
    if (r >= n) return;
    const int b = rp[r], e = rp[r + 1];
    offcnt[r] = 0;
    diag_pos[r] = -1;
    dis[r] = 1.0;
    if (b > e || b < 0 || static_cast<long long>(e) > nnz) {
        atomicOr(err, kErrRowPtr);
        return;
    }
    int dp = -1, cnt = 0;
    for (int o = b; o < e; ++o) {
        const int c = ci[o];
        if (c < 0 || c >= n) {
            atomicOr(err, kErrCol);
            return;
        }
        if (c == r && dp < 0)
            dp = o;
        else
            ++cnt;
    }
    offcnt[r] = cnt;
    diag_pos[r] = dp;
    if (dp >= 0 && va[dp] > 0.0) dis[r] = 1.0 / sqrt(va[dp]);
}

// t_r = sum over the row's off-diagonal entries of |a| d_r d_c, CSR order,
// each product rounded before it is added (as the host computes it).
__global__ void k_totals(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                         const double* __restrict__ va, const double* __restrict__ dis, double* __restrict__ t) {
#pragma clang fp contract(off)
    const int r = static_cast<int>(blockIdx.x) * kRowThreads + static_cast<int>(threadIdx.x);
    if (r >= n) return;
    const double dr = dis[r];
    double acc = 0.0;
    for (int o = rp[r]; o < rp[r + 1]; ++o) {
        const int c = ci[o];
        if (c == r) continue;
        const double sv = fabs(va[o]) * dr * dis[c];
        acc += sv;
    }
    t[r] = acc;
}

// ---- aggregation ----------------------------------------------------------------------
// Row a precedes row b in the seed order: larger total strength, then the
// smaller row (b < 0: none yet).
__device__ __forceinline__ bool precedes(int a, int b, const double* __restrict__ t) {
    if (b < 0) return true;
    const double ta = t[a], tb = t[b];
    return ta > tb || (ta == tb && a < b);
}

__global__ void k_seed_best1(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                             const double* __restrict__ t, const uint8_t* __restrict__ elig,
                             const int* __restrict__ owner, int* __restrict__ m1) {
    const int i = static_cast<int>(blockIdx.x) * kRowThreads + static_cast<int>(threadIdx.x);
    if (i >= n) return;
    int best = (elig[i] && owner[i] < 0) ? i : -1;
    for (int o = rp[i]; o < rp[i + 1]; ++o) {
        const int c = ci[o];
        if (c != i && elig[c] && owner[c] < 0 && precedes(c, best, t)) best = c;
    }
    m1[i] = best;
}

// Seeds: the free rows that precede every free row within distance 2
// (dist2), or within distance 1 (the variant for dense coarse graphs, where
// distance-2 neighbourhoods admit few seeds per round).
__global__ void k_seed_pick(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                            const double* __restrict__ t, const uint8_t* __restrict__ elig,
                            const int* __restrict__ m1, int* __restrict__ owner, int* __restrict__ size,
                            int* __restrict__ members, int dist2) {
    const int i = static_cast<int>(blockIdx.x) * kRowThreads + static_cast<int>(threadIdx.x);
    if (i >= n || !elig[i] || owner[i] >= 0) return;
    int best = m1[i];
    for (int o = rp[i]; dist2 && o < rp[i + 1]; ++o) {
        const int c = ci[o];
        if (c == i) continue;
        const int mc = m1[c];
        if (mc >= 0 && precedes(mc, best, t)) best = mc;
    }
    if (best == i) {
        owner[i] = i;
        size[i] = 1;
        members[static_cast<size_t>(i) * 8] = i;
    }
}

// One growth step, bidding half: row j scores every neighbouring block with
// room by the strength its entries put into the block (summed in CSR order
// from the block's first entry) over t_j, picks the best (ties: smaller
// block id) and posts (FP32 score, -j) to that block with an integer max.
// One pass over the row keeps the distinct blocks met so far, with their
// running sums, in LDS (kBidCap per thread): every block's sum takes its
// entries in CSR order, as the definition says.  A row that meets more than
// kBidCap blocks with room scans its entries once per block instead.
constexpr int kBidCap = 8;
__global__ __launch_bounds__(kRowThreads) void k_grow_bid(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                           const double* __restrict__ va, const double* __restrict__ dis,
                           const double* __restrict__ t, const uint8_t* __restrict__ elig,
                           const int* __restrict__ owner, const int* __restrict__ size,
                           unsigned long long* __restrict__ bid, const int* __restrict__ prev_taken) {
#pragma clang fp contract(off)
    __shared__ int lb[kBidCap * kRowThreads];
    __shared__ double ls[kBidCap * kRowThreads];
    // a step after one that took no row would find the same bids: nothing to do
    if (prev_taken != nullptr && *prev_taken == 0) return;
    const int tid = static_cast<int>(threadIdx.x);
    const int j = static_cast<int>(blockIdx.x) * kRowThreads + tid;
    if (j >= n || !elig[j] || owner[j] >= 0) return;
    const int b = rp[j], e = rp[j + 1];
    const double dj = dis[j], tj = t[j];
    int best_b = -1;
    double best_f = -1.0;
    int* blk = lb + tid;
    double* sum = ls + tid;
    int d = 0;
    bool over = false;
    for (int o = b; o < e; ++o) {
        const int c = ci[o];
        if (c == j) continue;
        const int B = owner[c];
        if (B < 0 || size[B] >= 8) continue;
        const double sv = fabs(va[o]) * dj * dis[c];
        int k = 0;
        while (k < d && blk[k * kRowThreads] != B) ++k;
        if (k < d) {
            sum[k * kRowThreads] += sv;
        } else if (d < kBidCap) {
            blk[d * kRowThreads] = B;
            sum[d * kRowThreads] = 0.0 + sv;
            ++d;
        } else {
            over = true;
            break;
        }
    }
    if (!over) {
        for (int k = 0; k < d; ++k) {
            const int B = blk[k * kRowThreads];
            const double f = tj > 0.0 ? sum[k * kRowThreads] / tj : 0.0;
            if (f > best_f || (f == best_f && B < best_b)) {
                best_b = B;
                best_f = f;
            }
        }
    } else {
        for (int o = b; o < e; ++o) {
            const int c = ci[o];
            if (c == j) continue;
            const int B = owner[c];
            if (B < 0 || size[B] >= 8) continue;
            bool seen = false;
            for (int o2 = b; o2 < o && !seen; ++o2) {
                const int c2 = ci[o2];
                seen = c2 != j && owner[c2] == B;
            }
            if (seen) continue;
            double acc = 0.0;
            for (int o2 = o; o2 < e; ++o2) {
                const int c2 = ci[o2];
                if (c2 == j || owner[c2] != B) continue;
                const double sv = fabs(va[o2]) * dj * dis[c2];
                acc += sv;
            }
            const double f = tj > 0.0 ? acc / tj : 0.0;
            if (f > best_f || (f == best_f && B < best_b)) {
                best_b = B;
                best_f = f;
            }
        }
    }
    if (best_b < 0) return;
    const unsigned long long key = (static_cast<unsigned long long>(__float_as_uint(static_cast<float>(best_f))) << 32) |
                                   static_cast<unsigned long long>(0xFFFFFFFFu - static_cast<unsigned>(j));
    atomicMax(bid + best_b, key);
}

// Taking half: every block with a bid adds the winning row.
__global__ void k_grow_take(int n, unsigned long long* __restrict__ bid, int* __restrict__ owner,
                            int* __restrict__ size, int* __restrict__ members, const int* __restrict__ prev_taken,
                            int* __restrict__ taken) {
    if (prev_taken != nullptr && *prev_taken == 0) return;
    const int B = static_cast<int>(blockIdx.x) * kRowThreads + static_cast<int>(threadIdx.x);
    if (B >= n) return;
    const unsigned long long k = bid[B];
    if (k == 0) return;
    *taken = 1;  // every writer stores the same value
    const int j = static_cast<int>(0xFFFFFFFFu - static_cast<unsigned>(k & 0xFFFFFFFFull));
    const int sz = size[B];
    members[static_cast<size_t>(B) * 8 + static_cast<size_t>(sz)] = j;
    size[B] = sz + 1;
    owner[j] = B;
    bid[B] = 0;
}

constexpr int kCountGroups = 256;
__global__ void k_count_free(int n, const uint8_t* __restrict__ elig, const int* __restrict__ owner,
                             int* __restrict__ counter) {
    int v = 0;
    for (int i = static_cast<int>(blockIdx.x) * kRowThreads + static_cast<int>(threadIdx.x); i < n;
         i += kCountGroups * kRowThreads)
        v += (elig[i] && owner[i] < 0) ? 1 : 0;
    group_count(v, counter);
}

// Regroup: rows of blocks with fewer than 8 rows become the eligible pool.
__global__ void k_pool(int n, int* __restrict__ owner, const int* __restrict__ size, uint8_t* __restrict__ elig) {
    const int r = static_cast<int>(blockIdx.x) * kRowThreads + static_cast<int>(threadIdx.x);
    if (r >= n) return;
    const bool pooled = size[owner[r]] < 8;
    elig[r] = pooled ? 1 : 0;
    if (pooled) owner[r] = -1;
}

__global__ void k_pool_sizes(int n, int* __restrict__ size) {
    const int B = static_cast<int>(blockIdx.x) * kRowThreads + static_cast<int>(threadIdx.x);
    if (B < n && size[B] < 8) size[B] = 0;
}

__device__ __forceinline__ void sort8(int (&v)[8], int cnt) {
    for (int a = 1; a < 8; ++a) {
        if (a >= cnt) break;
        const int x = v[a];
        int b = a - 1;
        while (b >= 0 && v[b] > x) {
            v[b + 1] = v[b];
            --b;
        }
        v[b + 1] = x;
    }
}

// Rows of each block ascending; full blocks are final units, smaller ones
// pieces; both are keyed by their smallest row (unique: a partition).
__global__ void k_unit_mark(int n, const int* __restrict__ size, int* __restrict__ members,
                            int* __restrict__ flag_unit, int* __restrict__ unit_src, int* __restrict__ flag_piece,
                            int* __restrict__ piece_src) {
    const int B = static_cast<int>(blockIdx.x) * kRowThreads + static_cast<int>(threadIdx.x);
    if (B >= n) return;
    const int sz = size[B];
    if (sz == 0) return;
    int v[8];
    int* m = members + static_cast<size_t>(B) * 8;
    for (int k = 0; k < 8; ++k) v[k] = k < sz ? m[k] : 0x7FFFFFFF;
    sort8(v, sz);
    for (int k = 0; k < sz; ++k) m[k] = v[k];
    if (sz == 8) {
        flag_unit[v[0]] = 1;
        unit_src[v[0]] = B;
    } else {
        flag_piece[v[0]] = 1;
        piece_src[v[0]] = B;
    }
}

__global__ void k_compact(int n, const int* __restrict__ flag, const int* __restrict__ pos,
                          const int* __restrict__ src, int* __restrict__ out) {
    const int r = static_cast<int>(blockIdx.x) * kRowThreads + static_cast<int>(threadIdx.x);
    if (r < n && flag[r]) out[pos[r]] = src[r];
}

__global__ void k_piece_sizes(int npieces, const int* __restrict__ pieces, const int* __restrict__ size,
                              int* __restrict__ psz) {
    const int p = static_cast<int>(blockIdx.x) * kRowThreads + static_cast<int>(threadIdx.x);
    if (p < npieces) psz[p] = size[pieces[p]];
}

// One thread per run of kPackSegment pieces (in smallest-row order): each
// piece into the most recent of the run's last kPackWindow packed blocks
// with room, else a new one.  Packed block k of run g is pid g * kPackSegment
// + k; pdst[p] = 8 pid + the piece's first slot.  The run's sizes are read
// up front, so the choice loop waits on no memory.
__global__ void k_pack(int npieces, const int* __restrict__ psz, int* __restrict__ pdst, int* __restrict__ psize) {
    const int g = static_cast<int>(blockIdx.x) * kRowThreads + static_cast<int>(threadIdx.x);
    const int p0 = g * kPackSegment;
    if (p0 >= npieces) return;
    const int cnt = min(npieces - p0, kPackSegment);
    int sz[kPackSegment], ps[kPackSegment];
#pragma unroll
    for (int k = 0; k < kPackSegment; ++k) {
        sz[k] = k < cnt ? psz[p0 + k] : 0;
        ps[k] = 0;
    }
    int np = 0;
    for (int k = 0; k < cnt; ++k) {
        int dst = -1;
        for (int q = np - 1; q >= max(0, np - kPackWindow); --q)
            if (ps[q] + sz[k] <= 8) {
                dst = q;
                break;
            }
        if (dst < 0) dst = np++;
        pdst[p0 + k] = (g * kPackSegment + dst) * 8 + ps[dst];
        ps[dst] += sz[k];
    }
    for (int k = 0; k < kPackSegment; ++k) psize[static_cast<size_t>(g) * kPackSegment + k] = ps[k];
}

__global__ void k_pack_copy(int npieces, const int* __restrict__ pieces, const int* __restrict__ psz,
                            const int* __restrict__ pdst, const int* __restrict__ members, int* __restrict__ prows) {
    const int p = static_cast<int>(blockIdx.x) * kRowThreads + static_cast<int>(threadIdx.x);
    if (p >= npieces) return;
    const int B = pieces[p];
    for (int i = 0; i < psz[p]; ++i) prows[pdst[p] + i] = members[static_cast<size_t>(B) * 8 + i];
}

__global__ void k_packed_mark(int npacked, int* __restrict__ prows, const int* __restrict__ psize,
                              int* __restrict__ flag_unit, int* __restrict__ unit_src) {
    const int pid = static_cast<int>(blockIdx.x) * kRowThreads + static_cast<int>(threadIdx.x);
    if (pid >= npacked) return;
    const int sz = psize[pid];
    if (sz == 0) return;
    int v[8];
    int* m = prows + static_cast<size_t>(pid) * 8;
    for (int k = 0; k < 8; ++k) v[k] = k < sz ? m[k] : 0x7FFFFFFF;
    sort8(v, sz);
    for (int k = 0; k < sz; ++k) m[k] = v[k];
    flag_unit[v[0]] = 1;
    unit_src[v[0]] = -(pid + 1);
}

__global__ void k_units_write(int n, const int* __restrict__ flag_unit, const int* __restrict__ upos,
                              const int* __restrict__ unit_src, const int* __restrict__ members,
                              const int* __restrict__ prows, const int* __restrict__ psize, int* __restrict__ slots) {
    const int r = static_cast<int>(blockIdx.x) * kRowThreads + static_cast<int>(threadIdx.x);
    if (r >= n || !flag_unit[r]) return;
    const int src = unit_src[r];
    const int* rows;
    int cnt;
    if (src >= 0) {
        rows = members + static_cast<size_t>(src) * 8;
        cnt = 8;
    } else {
        rows = prows + static_cast<size_t>(-src - 1) * 8;
        cnt = psize[-src - 1];
    }
    int* out = slots + static_cast<size_t>(upos[r]) * 8;
    for (int k = 0; k < 8; ++k) out[k] = k < cnt ? rows[k] : -1;
}

// ---- SELL build --------------------------------------------------------------------------
// Sort key of a unit: its longest row's off-diagonal count.
__global__ void k_unit_keys(int units, int br, const int* __restrict__ slots, const int* __restrict__ offcnt,
                            int* __restrict__ key) {
    const int u = static_cast<int>(blockIdx.x) * kRowThreads + static_cast<int>(threadIdx.x);
    if (u >= units) return;
    if (br == 1) {
        key[u] = offcnt[u];
        return;
    }
    int k = 0;
    for (int j = 0; j < 8; ++j) {
        const int r = slots[static_cast<size_t>(u) * 8 + j];
        if (r >= 0) k = max(k, offcnt[r]);
    }
    key[u] = k;
}

// Stable sort of each window of W units by decreasing key: bitonic sort of
// (complemented key, index) pairs in LDS, one workgroup per window.
__global__ __launch_bounds__(kSortThreads) void k_window_sort(int units, int W, int P, const int* __restrict__ key,
                                                              int* __restrict__ order) {
    extern __shared__ unsigned long long sk[];
    const int w0 = static_cast<int>(blockIdx.x) * W;
    const int cnt = min(W, units - w0);
    for (int i = static_cast<int>(threadIdx.x); i < P; i += kSortThreads)
        sk[i] = i < cnt ? (static_cast<unsigned long long>(0xFFFFFFFFu - static_cast<unsigned>(key[w0 + i])) << 32) |
                              static_cast<unsigned long long>(i)
                        : ~0ull;
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = static_cast<int>(threadIdx.x); i < P; i += kSortThreads) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const unsigned long long a = sk[i], b = sk[ixj];
                    const bool up = (i & k) == 0;
                    if ((a > b) == up) {
                        sk[i] = b;
                        sk[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int i = static_cast<int>(threadIdx.x); i < cnt; i += kSortThreads)
        order[w0 + i] = w0 + static_cast<int>(sk[i] & 0xFFFFFFFFull);
}

__global__ void k_perm(int units, int br, const int* __restrict__ order, const int* __restrict__ slots,
                       int* __restrict__ perm, int* __restrict__ iperm, int* __restrict__ block_unit) {
    const int pos = static_cast<int>(blockIdx.x) * kRowThreads + static_cast<int>(threadIdx.x);
    if (pos >= units) return;
    const int u = order[pos];
    if (br == 8) block_unit[pos] = u;
    for (int j = 0; j < br; ++j) {
        const int r = br == 8 ? slots[static_cast<size_t>(u) * 8 + j] : u;
        const int i = pos * br + j;
        perm[i] = r;
        if (r >= 0) iperm[r] = i;
    }
}

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}

// Slice shapes from the rows in natural order (coalesced CSR reads): every
// row raises its slice's width to 1 + its off-diagonal count and widens its
// slice's column span (in solve order) with integer atomics; spans start at
// the slice's own rows.  Then per slice: slice_base, the 16-bit test and
// the widest slice.
__global__ void k_slice_init(int nslices, int* __restrict__ width, int* __restrict__ smin, int* __restrict__ smax) {
    const int sl = static_cast<int>(blockIdx.x) * kRowThreads + static_cast<int>(threadIdx.x);
    if (sl >= nslices) return;
    width[sl] = 1;
    smin[sl] = sl * kC;
    smax[sl] = sl * kC + kC - 1;
}

__global__ void k_slice_rows(int n, const int* __restrict__ iperm, const int* __restrict__ rp,
                             const int* __restrict__ ci, const int* __restrict__ diag_pos,
                             const int* __restrict__ offcnt, int* __restrict__ width, int* __restrict__ smin,
                             int* __restrict__ smax) {
    const int r = static_cast<int>(blockIdx.x) * kRowThreads + static_cast<int>(threadIdx.x);
    if (r >= n) return;
    const int sl = iperm[r] / kC;
    const int dp = diag_pos[r];
    int cmin = 0x7FFFFFFF, cmax = -1;
    for (int o = rp[r]; o < rp[r + 1]; ++o) {
        if (o == dp) continue;
        const int c = iperm[ci[o]];
        cmin = min(cmin, c);
        cmax = max(cmax, c);
    }
    if (offcnt[r] > 0) atomicMax(width + sl, 1 + offcnt[r]);
    if (cmax >= 0) {
        atomicMin(smin + sl, cmin);
        atomicMax(smax + sl, cmax);
    }
}

__global__ void k_slice_finish(int nslices, const int* __restrict__ width, const int* __restrict__ smin,
                               const int* __restrict__ smax, int* __restrict__ slice_base, int* __restrict__ flags) {
    const int sl = static_cast<int>(blockIdx.x) * kRowThreads + static_cast<int>(threadIdx.x);
    int w = 0, wide = 0;
    if (sl < nslices) {
        slice_base[sl] = smin[sl];
        w = width[sl];
        wide = smax[sl] - smin[sl] > 65535 ? 1 : 0;
    }
    w = wave_max(w);
    wide = wave_max(wide);
    if ((threadIdx.x & 63) == 0) {
        if (wide) atomicAnd(flags, ~1);
        atomicMax(flags + 1, w);
    }
}

__device__ __forceinline__ int16_t lo16_dev(double v, float hi) {
    if (hi == 0.0f || !isfinite(hi)) return 0;
    int e = 0;
    (void)frexp(static_cast<double>(hi), &e);
    const double q = ldexp(v - static_cast<double>(hi), 38 - (e - 1));
    const double r = rint(q);
    return static_cast<int16_t>(fmax(-32767.0, fmin(32767.0, r)));
}

struct FillOut {
    float* val;
    float* val_lo;
    int16_t* val_lo16;
    double* val64;
    uint16_t* col16;
    int* col32;
};

__device__ __forceinline__ void put_entry(const FillOut& f, size_t e, double v, int col, int cb, bool diag_slot) {
    const float hi = static_cast<float>(v);
    f.val[e] = hi;
    if (f.val_lo != nullptr) {
        f.val_lo[e] = static_cast<float>(v - static_cast<double>(hi));
        f.val_lo16[e] = lo16_dev(v, hi);
    }
    if (f.val64 != nullptr) f.val64[e] = v;
    if (f.col16 != nullptr) f.col16[e] = diag_slot ? 0 : static_cast<uint16_t>(col - cb);
    if (f.col32 != nullptr) f.col32[e] = col;
}

// One thread per natural row: its slots in its slice (diagonal first, the
// off-diagonal entries in CSR order with columns in solve order, then
// padding slots pointing at the row itself with value 0).
__global__ void k_sell_fill_rows(int n, const int* __restrict__ iperm, const int* __restrict__ rp,
                                 const int* __restrict__ ci, const double* __restrict__ va,
                                 const int* __restrict__ diag_pos, const int* __restrict__ slice_ptr,
                                 const int* __restrict__ slice_base, FillOut f) {
    const int r = static_cast<int>(blockIdx.x) * kRowThreads + static_cast<int>(threadIdx.x);
    if (r >= n) return;
    const int i = iperm[r];
    const int sl = i / kC;
    const int cb = slice_base[sl];
    const int w = slice_ptr[sl + 1] - slice_ptr[sl];
    const size_t e0 = static_cast<size_t>(slice_ptr[sl]) * kC + static_cast<size_t>(i % kC);
    const int dp = diag_pos[r];
    put_entry(f, e0, dp >= 0 ? va[dp] : 0.0, i, cb, true);
    int j = 1;
    for (int o = rp[r]; o < rp[r + 1]; ++o) {
        if (o == dp) continue;
        put_entry(f, e0 + static_cast<size_t>(j) * kC, va[o], iperm[ci[o]], cb, false);
        ++j;
    }
    for (; j < w; ++j) put_entry(f, e0 + static_cast<size_t>(j) * kC, 0.0, i, cb, false);
}

// Padding rows (no natural row): identity.
__global__ void k_sell_fill_pad(int n_pad, const int* __restrict__ perm, const int* __restrict__ slice_ptr,
                                const int* __restrict__ slice_base, FillOut f) {
    const int i = static_cast<int>(blockIdx.x) * kRowThreads + static_cast<int>(threadIdx.x);
    if (i >= n_pad || perm[i] >= 0) return;
    const int sl = i / kC;
    const int cb = slice_base[sl];
    const int w = slice_ptr[sl + 1] - slice_ptr[sl];
    const size_t e0 = static_cast<size_t>(slice_ptr[sl]) * kC + static_cast<size_t>(i % kC);
    put_entry(f, e0, 1.0, i, cb, true);
    for (int j = 1; j < w; ++j) put_entry(f, e0 + static_cast<size_t>(j) * kC, 0.0, i, cb, false);
}

// Natural diagonal, solve-order 1 / diag, and the first row that fails the
// Jacobi checks (key 4 row + kind: the smallest key is the smallest row).
__global__ void k_diag_minv(int n_pad, int n, const int* __restrict__ perm, const int* __restrict__ diag_pos,
                            const double* __restrict__ va, double* __restrict__ diag, float* __restrict__ minv,
                            int* __restrict__ bad) {
    const int i = static_cast<int>(blockIdx.x) * kRowThreads + static_cast<int>(threadIdx.x);
    if (i >= n_pad) return;
    const int r = perm[i];
    if (r < 0) {
        minv[i] = 1.0f;
        return;
    }
    const int dp = diag_pos[r];
    const double d = dp >= 0 ? va[dp] : 0.0;
    diag[r] = d;
    minv[i] = static_cast<float>(1.0 / d);
    const int kind = dp < 0 ? 1 : (fabs(d) <= DBL_MIN ? 2 : (fabs(d) <= 1.0e-30 ? 3 : 0));
    if (kind != 0) atomicMin(bad, r * 4 + kind);
}

__global__ void k_natural_slots(int n, int units, int* __restrict__ slots) {
    const int i = static_cast<int>(blockIdx.x) * kRowThreads + static_cast<int>(threadIdx.x);
    if (i < units * 8) slots[i] = i < n ? i : -1;
}

constexpr int kSumGroups = kSysSumParts;  // workgroups per system

// Partial sums of x^2 over a system's rows: workgroup g of system k takes
// rows sys_begin[k] + g*256 + j*64*256 ..; then one fixed-order sum per system.
__global__ void k_sys_sumsq_part(const int* __restrict__ sys_begin, const double* __restrict__ x,
                                 double* __restrict__ part) {
    __shared__ double red[kRowThreads / 64];
    const int k = static_cast<int>(blockIdx.y), g = static_cast<int>(blockIdx.x);
    double acc = 0.0;
    for (int i = sys_begin[k] + g * kRowThreads + static_cast<int>(threadIdx.x); i < sys_begin[k + 1];
         i += kSumGroups * kRowThreads)
        acc += x[i] * x[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kRowThreads / 64; ++w) t += red[w];
        part[k * kSumGroups + g] = t;
    }
}

__global__ void k_sys_sumsq_finish(int nsys, const double* __restrict__ part, double* __restrict__ out) {
    const int k = static_cast<int>(blockIdx.x) * kRowThreads + static_cast<int>(threadIdx.x);
    if (k >= nsys) return;
    double t = 0.0;
    for (int g = 0; g < kSumGroups; ++g) t += part[k * kSumGroups + g];
    out[k] = t;
}

}  // namespace

int device_natural_slots(int n, DevBuf<int>& slots, hipStream_t s) {
    const int units = (n + 7) / 8;
    slots.alloc(static_cast<size_t>(units) * 8);
    hipLaunchKernelGGL(k_natural_slots, dim3(blocks_of(static_cast<long long>(units) * 8, kRowThreads)),
                       dim3(kRowThreads), 0, s, n, units, slots.p);
    HIPX(hipGetLastError());
    return units;
}

void launch_sys_sumsq(int nsys, const int* sys_begin, const double* x, double* out, double* part, hipStream_t s) {
    hipLaunchKernelGGL(k_sys_sumsq_part, dim3(kSumGroups, nsys), dim3(kRowThreads), 0, s, sys_begin, x, part);
    hipLaunchKernelGGL(k_sys_sumsq_finish, dim3(blocks_of(nsys, kRowThreads)), dim3(kRowThreads), 0, s, nsys, part,
                       out);
}

void device_scan_exclusive(const int* in, int* out, int n, int* total, hipStream_t s) {
    scan_exclusive(in, out, n, total, s);
    HIPX(hipGetLastError());
}

// ---- host side ------------------------------------------------------------------------------
std::mutex& set_up_mutex() {
    static std::mutex mu;
    return mu;
}

void device_rows(DevCsr& a, DevRows& rows, hipStream_t s, int trusted_nnz) {
    const int n = a.n;
    if (n <= 0) throw std::runtime_error("n must be positive");
    if (a.row_ptr == nullptr || a.col_idx == nullptr || a.values == nullptr)
        throw std::runtime_error("build_sell: null CSR array");
    if (trusted_nnz < 0) {
        int ends[2] = {0, 0};
        HIPX(hipMemcpyAsync(&ends[0], a.row_ptr, sizeof(int), hipMemcpyDeviceToHost, s));
        HIPX(hipMemcpyAsync(&ends[1], a.row_ptr + n, sizeof(int), hipMemcpyDeviceToHost, s));
        HIPX(hipStreamSynchronize(s));
        if (ends[0] != 0) throw std::runtime_error("row_ptr[0] must be 0");
        if (ends[1] < 0) throw std::runtime_error("row_ptr must be nondecreasing");
        a.nnz = ends[1];
    } else {
        a.nnz = trusted_nnz;
    }
    rows.offcnt = take<int>(kWOffcnt, static_cast<size_t>(n));
    rows.diag_pos = take<int>(kWDiagPos, static_cast<size_t>(n));
    rows.dis = take<double>(kWDis, static_cast<size_t>(n));
    rows.total = take<double>(kWTotal, static_cast<size_t>(n));
    int* err = take<int>(kWErr, 1);
    HIPX(hipMemsetAsync(err, 0, sizeof(int), s));
    const dim3 g(blocks_of(n, kRowThreads)), b(kRowThreads);
    hipLaunchKernelGGL(k_row_info, g, b, 0, s, n, a.nnz, a.row_ptr, a.col_idx, a.values, rows.offcnt,
                       rows.diag_pos, rows.dis, err);
    if (trusted_nnz < 0) {
        int e = 0;
        HIPX(hipMemcpyAsync(&e, err, sizeof(int), hipMemcpyDeviceToHost, s));
        HIPX(hipStreamSynchronize(s));
        if (e & kErrRowPtr) throw std::runtime_error("row_ptr must be nondecreasing");
        if (e & kErrCol) throw std::runtime_error("col_idx out of range");
    }
    hipLaunchKernelGGL(k_totals, g, b, 0, s, n, a.row_ptr, a.col_idx, a.values, rows.dis, rows.total);
    HIPX(hipGetLastError());
}

int device_aggregate(const DevCsr& a, const DevRows& rows, DevBuf<int>& slots, AggStats* st, hipStream_t s,
                     int seed_dist) {
    const int n = a.n;
    const size_t nn = static_cast<size_t>(n);
    int* owner = take<int>(kWOwner, nn);
    int* size = take<int>(kWSize, nn);
    int* members = take<int>(kWMembers, 8 * nn);
    int* m1 = take<int>(kWM1, nn);
    int* counter = take<int>(kWCounter, 1);
    int* taken = take<int>(kWTaken, kGrowSteps);
    unsigned long long* bid = take<unsigned long long>(kWBid, nn);
    uint8_t* elig = take<uint8_t>(kWElig, nn);
    fill(owner, n, -1, s);
    fill(size, n, 0, s);
    fill(bid, n, 0ull, s);
    fill(elig, n, static_cast<uint8_t>(1), s);
    const dim3 g(blocks_of(n, kRowThreads)), b(kRowThreads);
    AggStats local;
    auto rounds = [&](int& count) {
        for (;;) {
            hipLaunchKernelGGL(k_seed_best1, g, b, 0, s, n, a.row_ptr, a.col_idx, rows.total, elig, owner, m1);
            hipLaunchKernelGGL(k_seed_pick, g, b, 0, s, n, a.row_ptr, a.col_idx, rows.total, elig, m1, owner,
                               size, members, seed_dist == 1 ? 0 : 1);
            fill(taken, kGrowSteps, 0, s);
            for (int step = 0; step < kGrowSteps; ++step) {
                const int* prev = step > 0 ? taken + step - 1 : nullptr;
                hipLaunchKernelGGL(k_grow_bid, g, b, 0, s, n, a.row_ptr, a.col_idx, a.values, rows.dis,
                                   rows.total, elig, owner, size, bid, prev);
                hipLaunchKernelGGL(k_grow_take, g, b, 0, s, n, bid, owner, size, members, prev,
                                   taken + step);
            }
            ++count;
            HIPX(hipMemsetAsync(counter, 0, sizeof(int), s));
            hipLaunchKernelGGL(k_count_free, dim3(kCountGroups), b, 0, s, n, elig, owner, counter);
            int left = 0;
            HIPX(hipMemcpyAsync(&left, counter, sizeof(int), hipMemcpyDeviceToHost, s));
            HIPX(hipStreamSynchronize(s));
            if (left == 0) break;
        }
    };
    rounds(local.rounds);
    hipLaunchKernelGGL(k_pool, g, b, 0, s, n, owner, size, elig);
    hipLaunchKernelGGL(k_pool_sizes, g, b, 0, s, n, size);
    HIPX(hipMemsetAsync(counter, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_count_free, dim3(kCountGroups), b, 0, s, n, elig, owner, counter);
    int pooled = 0;
    HIPX(hipMemcpyAsync(&pooled, counter, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPX(hipStreamSynchronize(s));
    if (pooled > 0) rounds(local.regroup_rounds);

    // units and pieces, keyed by their smallest row
    int* flag_unit = take<int>(kWFlagUnit, nn);
    int* unit_src = take<int>(kWUnitSrc, nn);
    int* flag_piece = take<int>(kWFlagPiece, nn);
    int* piece_src = take<int>(kWPieceSrc, nn);
    int* pos = take<int>(kWPos, nn);
    int* totals = take<int>(kWTotals, 2);
    fill(flag_unit, n, 0, s);
    fill(flag_piece, n, 0, s);
    hipLaunchKernelGGL(k_unit_mark, g, b, 0, s, n, size, members, flag_unit, unit_src, flag_piece,
                       piece_src);
    scan_exclusive(flag_piece, pos, n, totals, s);
    int npieces = 0;
    HIPX(hipMemcpyAsync(&npieces, totals, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPX(hipStreamSynchronize(s));
    const int nseg = (npieces + kPackSegment - 1) / kPackSegment;
    const int npacked = nseg * kPackSegment;
    int* pieces = take<int>(kWPieces, static_cast<size_t>(npieces));
    int* psz = take<int>(kWPsz, static_cast<size_t>(npieces));
    int* pdst = take<int>(kWPdst, static_cast<size_t>(npieces));
    int* prows = take<int>(kWProws, static_cast<size_t>(npacked) * 8);
    int* psize = take<int>(kWPsize, static_cast<size_t>(npacked));
    if (npieces > 0) {
        const dim3 gp(blocks_of(npieces, kRowThreads));
        hipLaunchKernelGGL(k_compact, g, b, 0, s, n, flag_piece, pos, piece_src, pieces);
        hipLaunchKernelGGL(k_piece_sizes, gp, b, 0, s, npieces, pieces, size, psz);
        hipLaunchKernelGGL(k_pack, dim3(blocks_of(nseg, kRowThreads)), b, 0, s, npieces, psz, pdst, psize);
        hipLaunchKernelGGL(k_pack_copy, gp, b, 0, s, npieces, pieces, psz, pdst, members, prows);
        hipLaunchKernelGGL(k_packed_mark, dim3(blocks_of(npacked, kRowThreads)), b, 0, s, npacked, prows, psize,
                           flag_unit, unit_src);
    }
    scan_exclusive(flag_unit, pos, n, totals + 1, s);
    int units = 0;
    HIPX(hipMemcpyAsync(&units, totals + 1, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPX(hipStreamSynchronize(s));
    slots.alloc(static_cast<size_t>(units) * 8);
    hipLaunchKernelGGL(k_units_write, g, b, 0, s, n, flag_unit, pos, unit_src, members, prows, psize,
                       slots.p);
    HIPX(hipGetLastError());
    local.pieces = npieces;
    local.units = units;
    if (st != nullptr) *st = local;
    return units;
}

void device_sell(const DevCsr& a, const DevRows& rows, int sigma, bool keep_lo, bool keep64, int block_rows,
                 const int* d_slots, int units, DevSell& out, hipStream_t s) {
    if (sigma <= 0) sigma = kC;
    if (sigma % kC != 0) throw std::runtime_error("build_sell: sigma must be a multiple of 64");
    if (block_rows != 1 && block_rows != 8) throw std::runtime_error("build_sell: block_rows must be 1 or 8");
    const int br = block_rows;
    const int nu = br == 8 ? units : a.n;
    if (br == 8 && (d_slots == nullptr || units <= 0)) throw std::runtime_error("build_sell: blocks missing");
    const int W = sigma / br;
    if (W > 4096) throw std::runtime_error("device build: the sort window holds at most 4096 units");
    int P = 2;
    while (P < W) P <<= 1;
    const int n = a.n;
    out.n = n;
    out.nnz = a.nnz;
    out.sigma = sigma;
    out.units = nu;
    out.n_pad = std::max((nu * br + kC - 1) / kC * kC, kC);
    out.nslices = out.n_pad / kC;
    const dim3 b(kRowThreads);
    int* key = take<int>(kWKey, static_cast<size_t>(nu));
    int* order = take<int>(kWOrder, static_cast<size_t>(nu));
    int* width = take<int>(kWWidth, static_cast<size_t>(out.nslices));
    int* flags = take<int>(kWFlags, 4);
    hipLaunchKernelGGL(k_unit_keys, dim3(blocks_of(nu, kRowThreads)), b, 0, s, nu, br, d_slots, rows.offcnt, key);
    hipLaunchKernelGGL(k_window_sort, dim3(blocks_of(nu, W)), dim3(kSortThreads),
                       static_cast<size_t>(P) * sizeof(unsigned long long), s, nu, W, P, key, order);
    out.perm.alloc(static_cast<size_t>(out.n_pad));
    out.iperm.alloc(static_cast<size_t>(n));
    fill(out.perm.p, out.n_pad, -1, s);
    if (br == 8) {
        out.block_unit.alloc(static_cast<size_t>(out.n_pad / 8));
        fill(out.block_unit.p, out.n_pad / 8, -1, s);
    } else {
        out.block_unit.release();
    }
    hipLaunchKernelGGL(k_perm, dim3(blocks_of(nu, kRowThreads)), b, 0, s, nu, br, order, d_slots, out.perm.p,
                       out.iperm.p, out.block_unit.p);
    const int flags0[4] = {1, 0, 0x7FFFFFFF, 0};  // fits16, max width, bad key
    HIPX(hipMemcpyAsync(flags, flags0, sizeof flags0, hipMemcpyHostToDevice, s));
    out.slice_base.alloc(static_cast<size_t>(out.nslices));
    out.slice_ptr.alloc(static_cast<size_t>(out.nslices) + 1);
    int* smin = take<int>(kWSmin, static_cast<size_t>(out.nslices));
    int* smax = take<int>(kWSmax, static_cast<size_t>(out.nslices));
    const dim3 gsl(blocks_of(out.nslices, kRowThreads)), gr(blocks_of(n, kRowThreads));
    hipLaunchKernelGGL(k_slice_init, gsl, b, 0, s, out.nslices, width, smin, smax);
    hipLaunchKernelGGL(k_slice_rows, gr, b, 0, s, n, out.iperm.p, a.row_ptr, a.col_idx, rows.diag_pos,
                       rows.offcnt, width, smin, smax);
    hipLaunchKernelGGL(k_slice_finish, gsl, b, 0, s, out.nslices, width, smin, smax, out.slice_base.p, flags);
    scan_exclusive(width, out.slice_ptr.p, out.nslices, out.slice_ptr.p + out.nslices, s);
    out.diag.alloc(static_cast<size_t>(n));
    out.minv.alloc(static_cast<size_t>(out.n_pad));
    hipLaunchKernelGGL(k_diag_minv, dim3(blocks_of(out.n_pad, kRowThreads)), b, 0, s, out.n_pad, n, out.perm.p,
                       rows.diag_pos, a.values, out.diag.p, out.minv.p, flags + 2);
    int got[3] = {0, 0, 0}, cols = 0;
    HIPX(hipMemcpyAsync(got, flags, 3 * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPX(hipMemcpyAsync(&cols, out.slice_ptr.p + out.nslices, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPX(hipStreamSynchronize(s));
    out.idx16 = got[0] & 1;
    out.max_width = got[1];
    out.bad_row = got[2] == 0x7FFFFFFF ? -1 : got[2] / 4;
    out.bad_kind = got[2] == 0x7FFFFFFF ? 0 : got[2] % 4;
    out.stored = static_cast<long long>(cols) * kC;
    const size_t total = static_cast<size_t>(out.stored);
    out.val.alloc(total);
    out.val_lo.release();
    out.val_lo16.release();
    out.val64.release();
    out.col16.release();
    out.col32.release();
    if (keep_lo) {
        out.val_lo.alloc(total);
        out.val_lo16.alloc(total);
    }
    if (keep64) out.val64.alloc(total);
    if (out.idx16)
        out.col16.alloc(total);
    else
        out.col32.alloc(total);
    const FillOut f{out.val.p, out.val_lo.p, out.val_lo16.p, out.val64.p, out.col16.p, out.col32.p};
    hipLaunchKernelGGL(k_sell_fill_rows, gr, b, 0, s, n, out.iperm.p, a.row_ptr, a.col_idx, a.values,
                       rows.diag_pos, out.slice_ptr.p, out.slice_base.p, f);
    hipLaunchKernelGGL(k_sell_fill_pad, dim3(blocks_of(out.n_pad, kRowThreads)), b, 0, s, out.n_pad, out.perm.p,
                       out.slice_ptr.p, out.slice_base.p, f);
    HIPX(hipGetLastError());
}

}  // namespace xp

// kernels.hip -- hand-written gfx950 (CDNA4) kernels of the Jacobi-PCG hot path.
//
// Replaces the reference's five fused operations (archived/kernels/
// cg_kernels.cpp:43-158, HLS twins archived/kernels/*_kernel.cpp) and the
// on-chip PCG loop of kernels/cuper_pcg_control_kernel.cpp:1121-1222.
//
// Design points:
//  * SpMV is HBM-bound integer/gather work with no dense contraction, so no
//    MFMA: one 64-lane wavefront per 64-row SELL slice, one row per lane,
//    coalesced 4-byte value and 2-byte column streams, diagonal column first
//    (its column index is the row itself and is never stored or loaded).
//  * The pAp dot of a PCG step is fused into the SpMV epilogue (the x value
//    of the diagonal slot *is* p[row]); the init step r = b - Ax, z, p, rz, rr
//    is fused into the SpMV of x0 (as write_cuper_outputs<Init> does on the
//    FPGA, kernels/cuper_pcg_control_kernel.cpp:463-555).
//  * Reductions without atomics: FP64 products of the FP32 operands (exact),
//    summed per lane, per wavefront with cross-lane shuffles, per workgroup
//    through LDS, and written as one partial per workgroup.  The kernel that
//    needs the total (the next one in stream order) has every workgroup sum
//    the few hundred partials of its system itself, in a fixed order, so all
//    workgroups agree bit for bit and the result is reproducible run to run.
//    Kernel boundaries are the only synchronisation: no fences, no tickets
//    (a 2048-way atomic ticket cost ~45 us per kernel on MI355X).
//  * The scalar part of a PCG step (alpha, beta, convergence and breakdown
//    tests, trace) is evaluated redundantly by every workgroup from those
//    sums; the first workgroup of a system writes the next state slot (two
//    slots, alternating per iteration, so no kernel reads what it writes).
//    A whole iteration is three kernels with no host round trip and is
//    captured in a HIP graph.  A system that has stopped turns every later
//    kernel of its rows into a no-op (the state is carried forward), so the
//    iteration count stays exact when a graph runs past it.
//  * Workgroup -> rows mapping is XCD-aware: physical workgroup b runs on XCD
//    b % 8, so logical block (b % 8) * (grid / 8) + b / 8 gives each XCD one
//    contiguous range of rows and keeps the x entries they share in one L2.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "kernels.hpp"

namespace xp {

namespace {

constexpr double kBreakdownEps = 1.0e-30;  // include/cg_common.hpp:13 kBreakdownEps

__device__ __forceinline__ int logical_block() {
    const int g = static_cast<int>(gridDim.x);
    const int b = static_cast<int>(blockIdx.x);
    if (g % kNumXcd != 0) return b;
    return (b % kNumXcd) * (g / kNumXcd) + b / kNumXcd;
}

// The work item of logical block lb.  A single system uses the uniform
// split computed from the block index (no dependent load); batches read the
// work map.
__device__ __forceinline__ WorkItem work_item(const Launch& L, int lb) {
    if (L.u_blocks > 0) {
        WorkItem w;
        w.pad_ = 0;
        if (lb >= L.u_blocks) {
            w.sys = -1;
            w.slice_begin = w.slice_end = 0;
            return w;
        }
        w.sys = 0;
        w.slice_begin = static_cast<int>(static_cast<long long>(L.u_slices) * lb / L.u_blocks);
        w.slice_end = static_cast<int>(static_cast<long long>(L.u_slices) * (lb + 1) / L.u_blocks);
        return w;
    }
    return L.work[lb];
}

__device__ __forceinline__ int2 sys_range(const Launch& L, int sys) {
    if (L.u_blocks > 0) return make_int2(0, L.u_blocks);
    return L.sys_blocks[sys];
}

// v of the lane that DPP control CTRL names (both 32-bit halves moved).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = static_cast<unsigned>(
        __builtin_amdgcn_mov_dpp(static_cast<int>(static_cast<unsigned>(b)), CTRL, 0xF, 0xF, false));
    const unsigned hi = static_cast<unsigned>(
        __builtin_amdgcn_mov_dpp(static_cast<int>(static_cast<unsigned>(b >> 32)), CTRL, 0xF, 0xF, false));
    return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}

__device__ __forceinline__ double readlane_f64(double v, int lane) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<unsigned>(b)), lane));
    const unsigned hi =
        static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<unsigned>(b >> 32)), lane));
    return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}

// Sum of v over the 64 lanes, in every lane; all lanes must be active.
// Inside each 16-lane row by DPP moves (pairs, quads, the half-row mirror,
// the row mirror: each step adds the same two partial sums in every lane,
// so all lanes of a row hold the same bits), then the four row sums read
// with readlane and added as (row0 + row1) + (row2 + row3).  No LDS
// traffic and no waits on it, unlike a ds_bpermute butterfly.
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_f64<0xB1>(v);   // quad_perm(1, 0, 3, 2)
    v += dpp_f64<0x4E>(v);   // quad_perm(2, 3, 0, 1)
    v += dpp_f64<0x141>(v);  // row_half_mirror
    v += dpp_f64<0x140>(v);  // row_mirror
    return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}

// Sums NV values over the workgroup (kBlock threads).  Every thread gets the
// totals.  Fixed combination order: lanes by butterfly, then the kWaves
// wavefront sums in wave order by lanes 0..NV-1 of wave 0.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
    __shared__ double red[kWaves][NV];
    __shared__ double tot[NV];
    const int wave = static_cast<int>(threadIdx.x) >> 6;
    const int lane = static_cast<int>(threadIdx.x) & 63;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) red[wave][k] = v[k];
    }
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < NV) {
        double t = 0.0;
        for (int w = 0; w < kWaves; ++w) t += red[w][threadIdx.x];
        tot[threadIdx.x] = t;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = tot[k];
}

// Sum of NV partial slots over the logical blocks [range.x, range.y) of a
// system, written by an earlier kernel.  Every thread receives the totals;
// every workgroup computes them in the same order.
template <int NV>
__device__ __forceinline__ void sum_partials(const Launch& L, int sys, const int (&slots)[NV], double (&out)[NV]) {
    const int2 range = sys_range(L, sys);
#pragma unroll
    for (int k = 0; k < NV; ++k) out[k] = 0.0;
    for (int b = range.x + static_cast<int>(threadIdx.x); b < range.y; b += kBlock) {
#pragma unroll
        for (int k = 0; k < NV; ++k) out[k] += L.partials[static_cast<size_t>(slots[k]) * L.grid + b];
    }
    block_sum<NV>(out);
}

__device__ __forceinline__ void put_partial(const Launch& L, int slot, int lb, double v) {
    L.partials[static_cast<size_t>(slot) * L.grid + lb] = v;
}

// A fused kernel whose system does not iterate carries its workgroup's pAp
// partial into this launch's slot: a residual replacement may reactivate a
// converged system and rerun its last scalar step, which then reads the
// slot of whatever parity comes next.
__device__ __forceinline__ void carry_pap(const Launch& L, int lb) {
    if (threadIdx.x == 0)
        put_partial(L, L.pap_slot, lb, L.partials[static_cast<size_t>(L.pap_prev) * L.grid + lb]);
}

__device__ __forceinline__ bool bad(double v) { return !isfinite(v); }

// One SELL row: y = sum_j val * x[col] in FP32, diagonal first.  xdiag
// receives x[row].  Slices up to kFastW columns wide (every stencil matrix
// here) take a fully unrolled path that issues all value and column loads of
// the slice, then all x gathers, before the first FMA, so each wavefront has
// ~3 w loads in flight; the width test is wave-uniform (no divergence).
constexpr int kFastW = 10;

// Matrix streams read once per SpMV: V = 2 loads them non-temporally (they
// do not displace the vectors from the caches; faster when the matrix does
// not stay in the 256 MB MALL between SpMVs, slower when it does).
template <bool NT, typename T>
__device__ __forceinline__ T stream_load(const T* p) {
    if (NT) return __builtin_nontemporal_load(p);
    return *p;
}

template <bool IDX16, int V>
__device__ __forceinline__ float sell_row_f32(const SellView& A, int slice, int lane,
                                              const float* __restrict__ x, float& xdiag) {
    constexpr bool NT = V == 2;
    const int j0 = A.slice_ptr[slice];
    const int w = A.slice_ptr[slice + 1] - j0;
    const int row = slice * kC + lane;
    const size_t e0 = static_cast<size_t>(j0) * kC + static_cast<size_t>(lane);
    const int base = A.slice_base[slice];
    if (V >= 1 && w <= kFastW) {
        // Branch-free: slots past the row's width re-read its last slot (a
        // cached line) and contribute 0, so every load is issued before the
        // first use instead of one guarded load at a time.
        float v[kFastW], xv[kFastW];
        int c[kFastW];
#pragma unroll
        for (int k = 0; k < kFastW; ++k) {
            const size_t e = e0 + static_cast<size_t>(k < w ? k : w - 1) * kC;
            v[k] = k < w ? stream_load<NT>(A.val + e) : 0.0f;
            c[k] = k == 0 ? row
                          : (IDX16 ? base + static_cast<int>(stream_load<NT>(A.col16 + e)) : stream_load<NT>(A.col32 + e));
        }
#pragma unroll
        for (int k = 0; k < kFastW; ++k) xv[k] = x[c[k]];
        xdiag = xv[0];
        float acc = v[0] * xv[0];
#pragma unroll
        for (int k = 1; k < kFastW; ++k) acc = fmaf(v[k], xv[k], acc);
        return acc;
    }
    size_t e = e0;
    xdiag = x[row];
    float acc = A.val[e] * xdiag;
    int j = 1;
    // wide slices (the dense coarse levels of the multilevel hierarchy):
    // chunks of kChunk slots whose column loads, gathers and value loads are
    // all issued before the chunk's FMAs (the same order of summation)
    constexpr int kChunk = 8;
    for (; j + kChunk <= w; j += kChunk) {
        int c[kChunk];
        float v[kChunk], xv[kChunk];
#pragma unroll
        for (int k = 0; k < kChunk; ++k) {
            const size_t ek = e + static_cast<size_t>(k + 1) * kC;
            c[k] = IDX16 ? base + static_cast<int>(A.col16[ek]) : A.col32[ek];
            v[k] = A.val[ek];
        }
#pragma unroll
        for (int k = 0; k < kChunk; ++k) xv[k] = x[c[k]];
#pragma unroll
        for (int k = 0; k < kChunk; ++k) acc = fmaf(v[k], xv[k], acc);
        e += static_cast<size_t>(kChunk) * kC;
    }
#pragma unroll 4
    for (; j < w; ++j) {
        e += kC;
        const int c = IDX16 ? base + static_cast<int>(A.col16[e]) : A.col32[e];
        acc = fmaf(A.val[e], x[c], acc);
    }
    return acc;
}

template <bool IDX16, int VMODE>  // VMODE 0: val64; 1: val + val_lo
__device__ __forceinline__ double sell_row_f64(const SellView& A, int slice, int lane,
                                               const double* __restrict__ x) {
    const int j0 = A.slice_ptr[slice];
    const int j1 = A.slice_ptr[slice + 1];
    const int row = slice * kC + lane;
    size_t e = static_cast<size_t>(j0) * kC + static_cast<size_t>(lane);
    auto value = [&](size_t k) -> double {
        if (VMODE == 0) return A.val64[k];
        return static_cast<double>(A.val[k]) + static_cast<double>(A.val_lo[k]);
    };
    double acc = value(e) * x[row];
    const int base = A.slice_base[slice];
#pragma unroll 4
    for (int j = j0 + 1; j < j1; ++j) {
        e += kC;
        const int c = IDX16 ? base + static_cast<int>(A.col16[e]) : A.col32[e];
        acc = fma(value(e), x[c], acc);
    }
    return acc;
}

// A x for an FP32 vector x with FP64 values (VMODE 0: val64; 1: val +
// val_lo; 2: val + val_lo16, ~39 bits, 2 bytes less per entry) and FP64
// accumulation.  An FP64 value times an FP32 operand rounds
// once in FP64 and the row sum carries only FP64 rounding, so A x keeps its
// relative accuracy when the row's terms cancel (smooth x on an
// ill-conditioned matrix), which an FP32 sum does not.  xdiag = x[row].
template <bool IDX16, int VMODE, bool NT = false>
__device__ __forceinline__ double sell_row_acc(const SellView& A, int slice, int lane, const float* __restrict__ x,
                                               float& xdiag) {
    const int j0 = A.slice_ptr[slice];
    const int w = A.slice_ptr[slice + 1] - j0;
    const int row = slice * kC + lane;
    const size_t e0 = static_cast<size_t>(j0) * kC + static_cast<size_t>(lane);
    const int base = A.slice_base[slice];
    auto value = [&](size_t e) -> double {
        if (VMODE == 0) return stream_load<NT>(A.val64 + e);
        if (VMODE == 2) {  // hi + lo16 * 2^(E-38), the double's exponent field built from hi's
            const float hi = stream_load<NT>(A.val + e);
            const int eb = static_cast<int>((__float_as_uint(hi) >> 23) & 0xFFu);  // E + 127
            const double scale = __longlong_as_double(static_cast<long long>(eb - 127 - 38 + 1023) << 52);
            return static_cast<double>(hi) + static_cast<double>(stream_load<NT>(A.val_lo16 + e)) * scale;
        }
        return static_cast<double>(stream_load<NT>(A.val + e)) + static_cast<double>(stream_load<NT>(A.val_lo + e));
    };
    if (w <= kFastW) {
        // Branch-free, as in sell_row_f32: slots past the row's width re-read
        // its last slot and contribute 0.  Indices and gathers first, the
        // value pairs stream into the FP64 sum.
        float xv[kFastW];
        int c[kFastW];
#pragma unroll
        for (int k = 0; k < kFastW; ++k) {
            const size_t e = e0 + static_cast<size_t>(k < w ? k : w - 1) * kC;
            c[k] = k == 0 ? row
                          : (IDX16 ? base + static_cast<int>(stream_load<NT>(A.col16 + e)) : stream_load<NT>(A.col32 + e));
        }
#pragma unroll
        for (int k = 0; k < kFastW; ++k) xv[k] = x[c[k]];
        xdiag = xv[0];
        double acc = value(e0) * static_cast<double>(xv[0]);
#pragma unroll
        for (int k = 1; k < kFastW; ++k) {
            const double v = value(e0 + static_cast<size_t>(k < w ? k : w - 1) * kC);
            acc = fma(k < w ? v : 0.0, static_cast<double>(xv[k]), acc);
        }
        return acc;
    }
    size_t e = e0;
    xdiag = x[row];
    double acc = value(e) * static_cast<double>(xdiag);
#pragma unroll 4
    for (int j = 1; j < w; ++j) {
        e += kC;
        const int c = IDX16 ? base + static_cast<int>(A.col16[e]) : A.col32[e];
        acc = fma(value(e), static_cast<double>(x[c]), acc);
    }
    return acc;
}

// z = M r for the 8x8 block of this lane's row: the 8 rows of a block are 8
// consecutive lanes (blocks start at multiples of 8 in solve order).  With
// M = S C C^T S (BlockPre): u = S r, w = C^T u, z = S C w.  The factor is
// spread over the block's lanes: lane i loads only the 20 bytes of its row
// pair (row i of C) and its scale; it forms its terms C_ik u_i of every
// w_k, the 8 lanes sum them with a butterfly of DPP moves inside the 8-lane
// group (half-row mirror, then quad swaps), and each lane finishes its own
// z_i = s_i sum_k C_ik w_k.  Every lane of the wavefront must call it.
struct BlockRow {
    uint32_t w[5];  // the row pair's 10 FP16 entries
    float s;        // s_i
};

__device__ __forceinline__ BlockRow block_load(const BlockPre* __restrict__ M, int row) {
    const int i = row & 7;
    const uint32_t* c = reinterpret_cast<const uint32_t*>(M[row >> 3].c) + 5 * (i < 4 ? i : 7 - i);
    BlockRow o;
#pragma unroll
    for (int q = 0; q < 5; ++q) o.w[q] = c[q];
    o.s = M[row >> 3].s[i];
    return o;
}

template <int CTRL>
__device__ __forceinline__ float dpp_add(float v) {
    const int o = __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false);
    return v + __builtin_bit_cast(float, o);
}

__device__ __forceinline__ float block_apply_factor(const BlockRow& b, int i, float r) {
    // row i of C: the pair's halves off .. off + 7 (off = 0 for rows 0-3, 8 - i
    // for rows 4-7), entries past the diagonal masked to 0
    const int off = i < 4 ? 0 : 8 - i;
    const int dw = off >> 1;
    uint32_t e[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint32_t w0 = b.w[q], w1 = b.w[q + 1], w2 = q + 2 < 5 ? b.w[q + 2] : 0u, w3 = q + 3 < 5 ? b.w[q + 3] : 0u;
        const uint32_t lo = dw == 0 ? w0 : (dw == 1 ? w1 : w2);
        const uint32_t hi = dw == 0 ? w1 : (dw == 1 ? w2 : w3);
        e[q] = (off & 1) ? __builtin_amdgcn_alignbit(hi, lo, 16) : lo;
    }
    float c[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const unsigned short h = static_cast<unsigned short>((k & 1) ? (e[k >> 1] >> 16) : (e[k >> 1] & 0xFFFFu));
        c[k] = k <= i ? static_cast<float>(__builtin_bit_cast(_Float16, h)) : 0.0f;
    }
    const float u = b.s * r;
    float w[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) w[k] = c[k] * u;
    // sums over the 8 lanes of the block: pairs (i, 7 - i), then (i, i ^ 2), (i, i ^ 1)
#pragma unroll
    for (int k = 0; k < 8; ++k) w[k] = dpp_add<0x141>(w[k]);  // row_half_mirror
#pragma unroll
    for (int k = 0; k < 8; ++k) w[k] = dpp_add<0x4E>(w[k]);   // quad_perm(2, 3, 0, 1)
#pragma unroll
    for (int k = 0; k < 8; ++k) w[k] = dpp_add<0xB1>(w[k]);   // quad_perm(1, 0, 3, 2)
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) acc = fmaf(c[k], w[k], acc);
    return b.s * acc;
}

__device__ __forceinline__ float block_apply(const BlockPre* __restrict__ M, int row, float r) {
    const BlockRow b = block_load(M, row);
    return block_apply_factor(b, row & 7, r);
}

// Factors full FP32 blocks (solve order, block q = 64 floats row-major) into
// BlockPre: s_i = sqrt(M_ii), C = chol(S^-1 M S^-1) in FP64, rounded to FP16.
// A block that is not positive definite keeps its diagonal (C = I).  One
// thread per block.
__global__ void k_block_factor(int nblocks, const float* __restrict__ full, BlockPre* __restrict__ out) {
    const int q = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (q >= nblocks) return;
    const float* m = full + static_cast<size_t>(q) * 64;
    double s[8], a[8][8], l[8][8] = {};
    bool ok = true;
    for (int i = 0; i < 8; ++i) {
        const double d = static_cast<double>(m[i * 8 + i]);
        ok = ok && d > 0.0;
        s[i] = d > 0.0 ? sqrt(d) : 1.0;
    }
    for (int i = 0; i < 8; ++i)
        for (int k = 0; k < 8; ++k) a[i][k] = static_cast<double>(m[i * 8 + k]) / (s[i] * s[k]);
    for (int k = 0; k < 8 && ok; ++k) {
        double d = a[k][k];
        for (int t = 0; t < k; ++t) d -= l[k][t] * l[k][t];
        if (!(d > 0.0)) {
            ok = false;
            break;
        }
        l[k][k] = sqrt(d);
        for (int i = k + 1; i < 8; ++i) {
            double v = a[i][k];
            for (int t = 0; t < k; ++t) v -= l[i][t] * l[k][t];
            l[i][k] = v / l[k][k];
        }
    }
    BlockPre o;
    for (int t = 0; t < 40; ++t) o.c[t] = 0;
    for (int i = 0; i < 8; ++i)
        for (int k = 0; k <= i; ++k) {
            const float v = ok ? static_cast<float>(l[i][k]) : (i == k ? 1.0f : 0.0f);
            o.c[block_half_index(i, k)] = __builtin_bit_cast(unsigned short, static_cast<_Float16>(v));
        }
    for (int i = 0; i < 8; ++i) o.s[i] = static_cast<float>(s[i]);
    out[q] = o;
}

// The wavefront index is made visibly uniform (a scalar): loop bounds and
// slice metadata then live in scalar registers and branches on them are
// uniform.
#define XP_WAVE_LANE                                                                    \
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) >> 6); \
    const int lane = static_cast<int>(threadIdx.x) & 63;

// ---------------------------------------------------------------------------
// PCG init: r = b - A x, z = minv r, p = z; partials of rz, rr.  Mirrors
// init_pcg_kernel (archived/kernels/cg_kernels.cpp:73-103) fused with the
// x0 SpMV.
template <bool IDX16, int V>
__global__ __launch_bounds__(kBlock, 8) void k_pcg_init(Launch L, SellView A, const float* __restrict__ x,
                                                     const float* __restrict__ b,
                                                     const float* __restrict__ minv,
                                                     const BlockPre* __restrict__ blockM,
                                                     float* __restrict__ r, float* __restrict__ z,
                                                     float* __restrict__ p) {
    const int lb = logical_block();
    const WorkItem w = work_item(L, lb);
    if (w.sys < 0) return;
    XP_WAVE_LANE
    double v[2] = {0.0, 0.0};
    for (int sl = w.slice_begin + wave; sl < w.slice_end; sl += kWaves) {
        const int row = sl * kC + lane;
        float xd;
        const float ax = sell_row_f32<IDX16, V>(A, sl, lane, x, xd);
        const float rv = b[row] - ax;
        const float zv = blockM != nullptr ? block_apply(blockM, row, rv)
                                           : static_cast<float>(static_cast<double>(minv[row]) * static_cast<double>(rv));
        r[row] = rv;
        z[row] = zv;
        p[row] = zv;
        v[0] += static_cast<double>(rv) * static_cast<double>(zv);
        v[1] += static_cast<double>(rv) * static_cast<double>(rv);
    }
    block_sum<2>(v);
    if (threadIdx.x == 0) {
        put_partial(L, kSlotRz, lb, v[0]);
        put_partial(L, kSlotRr, lb, v[1]);
    }
}

// One workgroup per system: the state after init, with the loop-top tests of
// the on-chip PCG loop (kernels/cuper_pcg_control_kernel.cpp:1123-1130).
// S holds each system's configuration (tau, max_iters, ...) on entry.
__global__ __launch_bounds__(kBlock, 8) void k_pcg_init_state(Launch L, SysState* S) {
    const int sys = static_cast<int>(blockIdx.x);
    double t[2];
    sum_partials<2>(L, sys, {kSlotRz, kSlotRr}, t);
    if (threadIdx.x != 0) return;
    SysState& s = S[sys];
    s.rz = t[0];
    s.rr = t[1];
    s.rr0 = t[1];
    s.best_rr = t[1];
    s.stall = 0;
    s.iters = 0;
    s.pap = 0.0;
    s.alpha = 0.0;
    s.beta = 0.0;
    s.fresh = 1;
    s.x_pend = 0;
    if (s.rr <= s.tau) {
        s.status = kStatusConverged;
        s.active = 0;
    } else if (bad(s.rz) || bad(s.rr) || fabs(s.rz) <= kBreakdownEps) {
        s.status = kStatusBreakdown;
        s.active = 0;
    } else if (s.max_iters <= 0) {
        s.status = kStatusMaxIter;
        s.active = 0;
    } else {
        s.status = kStatusRunning;
        s.active = 1;
    }
}

// ap = A p and the partials of pAp.
template <bool IDX16, int V>
__global__ __launch_bounds__(kBlock, 8) void k_pcg_spmv(Launch L, SellView A, const SysState* S,
                                                     const float* __restrict__ p, float* __restrict__ ap) {
    const int lb = logical_block();
    const WorkItem w = work_item(L, lb);
    if (w.sys < 0 || !S[w.sys].active) return;
    XP_WAVE_LANE
    double v[1] = {0.0};
    for (int sl = w.slice_begin + wave; sl < w.slice_end; sl += kWaves) {
        float pd;
        const float y = sell_row_f32<IDX16, V>(A, sl, lane, p, pd);
        ap[sl * kC + lane] = y;
        v[0] += static_cast<double>(pd) * static_cast<double>(y);
    }
    block_sum<1>(v);
    if (threadIdx.x == 0) put_partial(L, L.pap_slot, lb, v[0]);
}

// Rows handled per thread in one pass of the vector kernels: every load of a
// pass is issued before the reduction prologue, so the prologue's latency
// (summing the partials of the previous kernel) overlaps the HBM reads.
constexpr int kRowsPerThread = 4;

// alpha = rz / pAp; x += alpha p; r -= alpha ap; z = minv r; partials of
// rz_new, rr_new.  update_xrz_kernel (archived/kernels/cg_kernels.cpp:118-147).
// A breakdown (bad pAp or alpha) leaves the vectors untouched; the next
// kernel records it.
template <bool BLOCK>
__global__ __launch_bounds__(kBlock, 8) void k_pcg_update_xrz(Launch L, const SysState* S, float* __restrict__ x,
                                                           const float* __restrict__ p,
                                                           float* __restrict__ r,
                                                           const float* __restrict__ ap,
                                                           const float* __restrict__ minv,
                                                           const BlockPre* __restrict__ blockM,
                                                           float* __restrict__ z) {
    // with blocks one row per pass (two or more measured slower in the
    // accurate form's update, below)
    constexpr int ROWS = BLOCK ? 1 : kRowsPerThread;
    const int lb = logical_block();
    const WorkItem w = work_item(L, lb);
    if (w.sys < 0) return;
    // Everything the first pass needs is requested at once (state, vectors,
    // partials); the activity test waits for all of it in one round trip.
    const int active = S[w.sys].active;
    const double rz_old = S[w.sys].rz;
    const int r0 = w.slice_begin * kC, r1 = w.slice_end * kC;
    double v[2] = {0.0, 0.0};
    bool first = true;
    double alpha = 0.0;
    for (int base = r0; base < r1; base += ROWS * kBlock) {
        float xv[ROWS], pv[ROWS], rv[ROWS], av[ROWS], mv[ROWS];
        BlockRow bl[ROWS];
#pragma unroll
        for (int k = 0; k < ROWS; ++k) {
            const int row = base + k * kBlock + static_cast<int>(threadIdx.x);
            if (row < r1) {
                xv[k] = x[row];
                pv[k] = p[row];
                rv[k] = r[row];
                av[k] = ap[row];
                mv[k] = BLOCK ? 0.0f : minv[row];
                if (BLOCK) bl[k] = block_load(blockM, row);
            }
        }
        if (first) {
            first = false;
            double t[1];
            sum_partials<1>(L, w.sys, {L.pap_slot}, t);
            if (!active) return;
            alpha = rz_old / t[0];
            if (bad(t[0]) || fabs(t[0]) <= kBreakdownEps || bad(alpha)) return;
        }
#pragma unroll
        for (int k = 0; k < ROWS; ++k) {
            const int row = base + k * kBlock + static_cast<int>(threadIdx.x);
            if (row < r1) {  // uniform per wavefront: r1 - base is a multiple of 64
                const float xn = static_cast<float>(static_cast<double>(xv[k]) + alpha * static_cast<double>(pv[k]));
                const float rn = static_cast<float>(static_cast<double>(rv[k]) - alpha * static_cast<double>(av[k]));
                const float zn = BLOCK
                                     ? block_apply_factor(bl[k], row & 7, rn)
                                     : static_cast<float>(static_cast<double>(mv[k]) * static_cast<double>(rn));
                x[row] = xn;
                r[row] = rn;
                z[row] = zn;
                v[0] += static_cast<double>(rn) * static_cast<double>(zn);
                v[1] += static_cast<double>(rn) * static_cast<double>(rn);
            }
        }
    }
    block_sum<2>(v);
    if (threadIdx.x == 0) {
        put_partial(L, kSlotRz, lb, v[0]);
        put_partial(L, kSlotRr, lb, v[1]);
    }
}

// Evaluates the scalar tail of an iteration from the partials: the state
// after the iteration, in `s`; identical in every thread of every workgroup
// of the system.  Golden loop src/CgSolverGolden.hpp:100-118; breakdown
// tests of archived/host/multi_kernel_solver.hpp:170-203; order of the
// post-update tests as the on-chip loop (kernels/cuper_pcg_control_kernel.cpp:
// 1193-1214), then the loop bound and the next loop-top rz test (:1127-1130).
__device__ __forceinline__ void pcg_scalar_step(SysState& s, const double (&t)[3], TraceRow* trace_row) {
    const double pap = t[0];
    const double alpha = s.rz / pap;
    s.pap = pap;
    if (bad(pap) || fabs(pap) <= kBreakdownEps || bad(alpha)) {
        s.status = kStatusBreakdown;
        s.active = 0;
        return;
    }
    const double rz_old = s.rz;
    const double rz = t[1];
    const double rr = t[2];
    const double beta = rz / rz_old;
    if (trace_row != nullptr) {
        trace_row->alpha = alpha;
        trace_row->beta = beta;
        trace_row->rz = rz;
        trace_row->rr = rr;
    }
    s.iters += 1;
    s.rz_prev = rz_old;
    s.rz = rz;
    s.rr = rr;
    s.alpha = alpha;
    s.beta = beta;
    if (bad(rz) || bad(rr)) {
        s.status = kStatusBreakdown;
        s.active = 0;
    } else if (rr <= s.tau) {
        s.status = kStatusConverged;
        s.active = 0;
    } else if (fabs(rz_old) <= kBreakdownEps || bad(beta)) {
        s.status = kStatusBreakdown;
        s.active = 0;
    } else if (s.iters >= s.max_iters) {
        s.status = kStatusMaxIter;
        s.active = 0;
    } else if (fabs(rz) <= kBreakdownEps) {
        s.status = kStatusBreakdown;
        s.active = 0;
    } else if (s.stall_limit > 0) {
        // Stagnation watch (for preconditioners that may fail): stop when
        // r^T r has not dropped below half its best for stall_limit iterations.
        if (rr < 0.5 * s.best_rr) {
            s.best_rr = rr;
            s.stall = 0;
        } else if (++s.stall >= s.stall_limit) {
            s.status = kStatusStagnation;
            s.active = 0;
        }
    }
}

// The scalar tail of the iteration and p = z + beta p
// (update_p_kernel, archived/kernels/cg_kernels.cpp:149-158).  Reads state
// slot Sc, writes slot Sn.
__global__ __launch_bounds__(kBlock, 8) void k_pcg_update_p(Launch L, const SysState* Sc, SysState* Sn,
                                                            TraceRow* trace, int trace_stride,
                                                            const float* __restrict__ z, float* __restrict__ p) {
    __shared__ SysState s_sh;  // the scalar step runs on LDS: no registers held across it
    const int lb = logical_block();
    const WorkItem w = work_item(L, lb);
    if (w.sys < 0) return;
    const bool leader_block = lb == sys_range(L, w.sys).x;
    if (threadIdx.x == 0) s_sh = Sc[w.sys];  // in flight with the loads below
    const int r0 = w.slice_begin * kC, r1 = w.slice_end * kC;
    bool first = true;
    for (int base = r0; base < r1; base += kRowsPerThread * kBlock) {
        float zv[kRowsPerThread], pv[kRowsPerThread];
#pragma unroll
        for (int k = 0; k < kRowsPerThread; ++k) {
            const int row = base + k * kBlock + static_cast<int>(threadIdx.x);
            if (row < r1) {
                zv[k] = z[row];
                pv[k] = p[row];
            }
        }
        if (first) {
            first = false;
            double t[3];
            sum_partials<3>(L, w.sys, {L.pap_slot, kSlotRz, kSlotRr}, t);  // ends with a barrier
            if (!s_sh.active) {
                if (leader_block && threadIdx.x == 0) Sn[w.sys] = s_sh;  // carry the stopped state forward
                return;
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                TraceRow* tr = nullptr;
                if (leader_block && trace != nullptr && s_sh.iters < s_sh.trace_cap)
                    tr = trace + static_cast<size_t>(w.sys) * static_cast<size_t>(trace_stride) +
                         static_cast<size_t>(s_sh.iters);
                pcg_scalar_step(s_sh, t, tr);
                if (leader_block) Sn[w.sys] = s_sh;
            }
            __syncthreads();
            if (!s_sh.active) return;
        }
        const double beta = s_sh.beta;
#pragma unroll
        for (int k = 0; k < kRowsPerThread; ++k) {
            const int row = base + k * kBlock + static_cast<int>(threadIdx.x);
            if (row < r1) p[row] = static_cast<float>(static_cast<double>(zv[k]) + beta * static_cast<double>(pv[k]));
        }
    }
}

// Fused iteration, first kernel: finishes the previous iteration (scalar
// tail from the rz/rr partials, as k_pcg_update_p does) and forms the new
// direction without a separate pass, using A p_new = A z + beta A p_old:
//   p = z + beta p;  ap = A z + beta ap;  partials of pAp.
// Right after init or a restart (state.fresh) p = z and ap = A z.  Reads slot
// Sc, writes slot Sn; the k_pcg_update_xrz that follows reads Sn.
template <bool IDX16, int V>
__global__ __launch_bounds__(kBlock, 8) void k_pcg_spmv_fused(Launch L, SellView A, const SysState* Sc,
                                                              SysState* Sn, TraceRow* trace, int trace_stride,
                                                              const float* __restrict__ z, float* __restrict__ p,
                                                              float* __restrict__ ap) {
    __shared__ SysState s_sh;
    const int lb = logical_block();
    const WorkItem w = work_item(L, lb);
    if (w.sys < 0) return;
    const bool leader_block = lb == sys_range(L, w.sys).x;
    // Prologue first (state and partials requested together), then a
    // barrier-free streaming loop.
    if (threadIdx.x == 0) s_sh = Sc[w.sys];
    double t[3];
    sum_partials<3>(L, w.sys, {L.pap_prev, kSlotRz, kSlotRr}, t);  // ends with a barrier
    if (!s_sh.active) {
        if (leader_block && threadIdx.x == 0) Sn[w.sys] = s_sh;
        carry_pap(L, lb);
        return;
    }
    const bool fresh = s_sh.fresh != 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        if (!fresh) {
            TraceRow* tr = nullptr;
            if (leader_block && trace != nullptr && s_sh.iters < s_sh.trace_cap)
                tr = trace + static_cast<size_t>(w.sys) * static_cast<size_t>(trace_stride) +
                     static_cast<size_t>(s_sh.iters);
            pcg_scalar_step(s_sh, t, tr);
        } else {
            s_sh.fresh = 0;
        }
        if (leader_block) Sn[w.sys] = s_sh;
    }
    __syncthreads();
    if (!s_sh.active) {
        carry_pap(L, lb);
        return;
    }
    const double beta = s_sh.beta;
    XP_WAVE_LANE
    double v[1] = {0.0};
    for (int sl = w.slice_begin + wave; sl < w.slice_end; sl += kWaves) {
        const int row = sl * kC + lane;
        float zd;
        const float az = sell_row_f32<IDX16, V>(A, sl, lane, z, zd);
        float pn, an;
        if (fresh) {  // after init or a restart from a replaced residual: p = z, ap = A z
            pn = zd;
            an = az;
        } else {
            pn = static_cast<float>(static_cast<double>(zd) + beta * static_cast<double>(p[row]));
            an = static_cast<float>(static_cast<double>(az) + beta * static_cast<double>(ap[row]));
        }
        p[row] = pn;
        ap[row] = an;
        v[0] += static_cast<double>(pn) * static_cast<double>(an);
    }
    block_sum<1>(v);
    if (threadIdx.x == 0) put_partial(L, L.pap_slot, lb, v[0]);
}

// ---------------------------------------------------------------------------
// Accurate iteration (FP32 r and z, FP64 p, A p and x; FP64-accumulated SpMV
// with hi+lo or FP64 values).  Same state machine, partial slots and scalar
// step as the FP32 kernels above; only the vector arithmetic differs.  With
// FP64 p and A p the recurrence A p = A z + beta A p_old stays consistent
// with p, so the recursive residual tracks b - A x.

// First kernel: scalar tail of the previous iteration, p = z + beta p,
// ap = A z + beta ap, pAp partials.  Right after init (fresh): p = z, ap = A z.
// A z does not depend on beta, so every wavefront computes its first slice's
// A z (and loads its p, ap) before waiting on the cross-workgroup partial
// sums and the scalar step: with ~2-3 slices per wavefront that wait would
// otherwise be exposed.
template <bool IDX16, int VMODE, bool NT>
__global__ __launch_bounds__(kBlock, 8) void k_pcg_spmv_fused_acc(Launch L, SellView A, const SysState* Sc,
                                                                  SysState* Sn, TraceRow* trace, int trace_stride,
                                                                  const float* __restrict__ z,
                                                                  const double* __restrict__ p_old,
                                                                  double* __restrict__ p_new,
                                                                  double* __restrict__ ap) {
    __shared__ SysState s_sh;
    const int lb = logical_block();
    const WorkItem w = work_item(L, lb);
    if (w.sys < 0) return;
    const bool leader_block = lb == sys_range(L, w.sys).x;
    if (threadIdx.x == 0) s_sh = Sc[w.sys];
    XP_WAVE_LANE
    // the first slice, ahead of the scalar step (right after init p and ap
    // hold stale values; they are loaded but not used then)
    int sl = w.slice_begin + wave;
    const bool have0 = sl < w.slice_end;
    double po0 = 0.0, apo0 = 0.0, az0 = 0.0;
    float zd0 = 0.0f;
    if (have0) {
        const int row = sl * kC + lane;
        po0 = p_old[row];
        apo0 = ap[row];
        az0 = sell_row_acc<IDX16, VMODE, NT>(A, sl, lane, z, zd0);
    }
    double t[3];
    sum_partials<3>(L, w.sys, {L.pap_prev, kSlotRz, kSlotRr}, t);  // ends with a barrier
    if (!s_sh.active) {
        if (leader_block && threadIdx.x == 0) Sn[w.sys] = s_sh;
        carry_pap(L, lb);
        return;
    }
    const bool fresh = s_sh.fresh != 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        if (!fresh) {
            TraceRow* tr = nullptr;
            if (leader_block && trace != nullptr && s_sh.iters < s_sh.trace_cap)
                tr = trace + static_cast<size_t>(w.sys) * static_cast<size_t>(trace_stride) +
                     static_cast<size_t>(s_sh.iters);
            const double pap = t[0];
            pcg_scalar_step(s_sh, t, tr);
            // Stopping here after an even update (this launch is odd) leaves
            // that update's alpha p out of x, unless its alpha was unusable
            // (then the update made no step either).
            if (!s_sh.active && L.pap_slot == kSlotPap1 &&
                !(bad(pap) || fabs(pap) <= kBreakdownEps || bad(s_sh.alpha)))
                s_sh.x_pend = 1;
        } else {
            s_sh.fresh = 0;
        }
        if (leader_block) Sn[w.sys] = s_sh;
    }
    __syncthreads();
    if (!s_sh.active) {
        // both p buffers keep the last direction (a replacement may
        // reactivate the system at either parity)
        carry_pap(L, lb);
        for (int s2 = w.slice_begin + wave; s2 < w.slice_end; s2 += kWaves) {
            const int row = s2 * kC + lane;
            p_new[row] = p_old[row];
        }
        return;
    }
    const double beta = s_sh.beta;
    double v[1] = {0.0};
    if (have0) {
        const int row = sl * kC + lane;
        const double pn = fresh ? static_cast<double>(zd0) : static_cast<double>(zd0) + beta * po0;
        const double an = fresh ? az0 : az0 + beta * apo0;
        p_new[row] = pn;
        ap[row] = an;
        v[0] += pn * an;
    }
    for (sl += kWaves; sl < w.slice_end; sl += kWaves) {
        const int row = sl * kC + lane;
        // the FP64 p and A p streams are requested before the row's gathers
        const double po = fresh ? 0.0 : p_old[row];
        const double apo = fresh ? 0.0 : ap[row];
        float zd;
        const double az = sell_row_acc<IDX16, VMODE, NT>(A, sl, lane, z, zd);
        double pn, an;
        if (fresh) {
            pn = static_cast<double>(zd);
            an = az;
        } else {
            pn = static_cast<double>(zd) + beta * po;
            an = az + beta * apo;
        }
        p_new[row] = pn;
        ap[row] = an;
        v[0] += pn * an;
    }
    block_sum<1>(v);
    if (threadIdx.x == 0) put_partial(L, L.pap_slot, lb, v[0]);
}

// alpha = rz / pAp; r = fp32(r - alpha ap); z = M r; partials of rz_new,
// rr_new; x in FP64 every second iteration.  XM = 0 (even iteration): x is
// left alone; XM = 1 (odd): x += alpha_prev p_prev + alpha p, with alpha_prev
// the previous update's (the state's, from the scalar step that closed it)
// and p_prev its direction, still in the other p buffer: two steps' x
// traffic (x read and written, two directions) instead of four.  A system
// that stops after an even update keeps x_pend (k_pcg_flush_x_acc).  Rows in
// flight per thread: the FP64 streams double the registers per row.
constexpr int kRowsAcc = 2;
template <bool BLOCK, int XM>
__global__ __launch_bounds__(kBlock, 8) void k_pcg_update_xrz_acc(Launch L, const SysState* S,
                                                               double* __restrict__ x,
                                                               const double* __restrict__ p,
                                                               const double* __restrict__ p_prev,
                                                               float* __restrict__ r,
                                                               const double* __restrict__ ap,
                                                               const float* __restrict__ minv,
                                                               const BlockPre* __restrict__ blockM,
                                                               float* __restrict__ z) {
    // One row per pass with blocks.  Measured on the 1.2M-row FEM matrix
    // (average of an even and an odd update, 256 / 512 workgroups): the
    // distributed block factor 12.50 / 12.33 us against 13.14 / 12.68 with
    // every lane loading the whole factor; two rows per pass 12.84 / 13.13,
    // three in the even launches 13.41 / 13.72; a software pipeline keeping
    // the loads of the next two passes in flight (three pass buffers in
    // fixed roles) 24.0 us at 512 workgroups.
    constexpr int ROWS = BLOCK ? 1 : kRowsAcc;
    const int lb = logical_block();
    const WorkItem w = work_item(L, lb);
    if (w.sys < 0) return;
    const int active = S[w.sys].active;
    const double rz_old = S[w.sys].rz;
    const double a_prev = XM ? S[w.sys].alpha : 0.0;
    const int r0 = w.slice_begin * kC, r1 = w.slice_end * kC;
    double v[2] = {0.0, 0.0};
    bool first = true;
    double alpha = 0.0;
    for (int base = r0; base < r1; base += ROWS * kBlock) {
        double xv[ROWS], pv[ROWS], qv[ROWS], av[ROWS];
        float rv[ROWS], mv[ROWS];
        BlockRow bl[ROWS];
#pragma unroll
        for (int k = 0; k < ROWS; ++k) {
            const int row = base + k * kBlock + static_cast<int>(threadIdx.x);
            if (row < r1) {
                if (XM) {
                    xv[k] = x[row];
                    pv[k] = p[row];
                    qv[k] = p_prev[row];
                }
                rv[k] = r[row];
                av[k] = ap[row];
                mv[k] = BLOCK ? 0.0f : minv[row];
                if (BLOCK) bl[k] = block_load(blockM, row);
            }
        }
        if (first) {
            first = false;
            double t[1];
            sum_partials<1>(L, w.sys, {L.pap_slot}, t);
            if (!active) return;
            alpha = rz_old / t[0];
            if (bad(t[0]) || fabs(t[0]) <= kBreakdownEps || bad(alpha)) {
                // no step here; the previous one still goes into x
                if (XM)
                    for (int row = r0 + static_cast<int>(threadIdx.x); row < r1; row += kBlock)
                        x[row] += a_prev * p_prev[row];
                return;
            }
        }
#pragma unroll
        for (int k = 0; k < ROWS; ++k) {
            const int row = base + k * kBlock + static_cast<int>(threadIdx.x);
            if (row < r1) {  // uniform per wavefront: r1 - base is a multiple of 64
                const float rn = static_cast<float>(static_cast<double>(rv[k]) - alpha * av[k]);
                const float zn = BLOCK
                                     ? block_apply_factor(bl[k], row & 7, rn)
                                     : static_cast<float>(static_cast<double>(mv[k]) * static_cast<double>(rn));
                if (XM) x[row] = (xv[k] + a_prev * qv[k]) + alpha * pv[k];
                r[row] = rn;
                z[row] = zn;
                v[0] += static_cast<double>(rn) * static_cast<double>(zn);
                v[1] += static_cast<double>(rn) * static_cast<double>(rn);
            }
        }
    }
    block_sum<2>(v);
    if (threadIdx.x == 0) {
        put_partial(L, kSlotRz, lb, v[0]);
        put_partial(L, kSlotRr, lb, v[1]);
    }
}

// The x update a system still lacks after stopping behind an even update
// (state.x_pend): x += alpha p with its last step's alpha and direction
// (both p buffers hold it).  Then k_pcg_clear_x_pend, a kernel of its own so
// that every workgroup here sees the flag set.
__global__ __launch_bounds__(kBlock, 8) void k_pcg_flush_x_acc(Launch L, const SysState* S,
                                                            const double* __restrict__ p, double* __restrict__ x) {
    const WorkItem w = work_item(L, logical_block());
    if (w.sys < 0) return;
    if (!S[w.sys].x_pend) return;
    const double alpha = S[w.sys].alpha;
    for (int row = w.slice_begin * kC + static_cast<int>(threadIdx.x); row < w.slice_end * kC; row += kBlock)
        x[row] += alpha * p[row];
}

__global__ void k_pcg_clear_x_pend(int nsys, SysState* S) {
    const int sys = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (sys < nsys) S[sys].x_pend = 0;
}

template <bool IDX16, int V>
__global__ __launch_bounds__(kBlock, 8) void k_spmv_f32(Launch L, SellView A, const float* __restrict__ x,
                                                     float* __restrict__ y) {
    const WorkItem w = work_item(L, logical_block());
    XP_WAVE_LANE
    for (int sl = w.slice_begin + wave; sl < w.slice_end; sl += kWaves) {
        float xd;
        y[sl * kC + lane] = sell_row_f32<IDX16, V>(A, sl, lane, x, xd);
    }
}

// Streaming reference kernels for the achievable HBM rate: every byte of a
// buffer read once with 16-byte loads (XCD-aware contiguous ranges, the
// SpMV's mapping), folded into one word per workgroup so the loads are not
// dropped; and a 16-byte copy (read + write).
__global__ __launch_bounds__(kBlock, 8) void k_stream_read(const uint4* __restrict__ src, size_t n16,
                                                        unsigned* __restrict__ sink) {
    const int lb = logical_block();
    const size_t per = (n16 + gridDim.x - 1) / gridDim.x;
    const size_t i0 = static_cast<size_t>(lb) * per, i1 = min(n16, i0 + per);
    unsigned acc = 0;
    for (size_t i = i0 + threadIdx.x; i < i1; i += kBlock) {
        const uint4 v = src[i];
        acc ^= v.x ^ v.y ^ v.z ^ v.w;
    }
    if (acc == 0x9E3779B9u) sink[lb] = acc;  // data-dependent, almost never taken
}

// The same read with non-temporal 16-byte loads, four in flight per thread
// (on this GPU the fastest way to stream a buffer once: 6.7-6.8 TB/s
// against 5.8-6.2 for the plain loads).
typedef unsigned stream4_t __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(kBlock, 8) void k_stream_read_nt(const uint4* __restrict__ src, size_t n16,
                                                           unsigned* __restrict__ sink) {
    const int lb = logical_block();
    const size_t per = (n16 + gridDim.x - 1) / gridDim.x;
    const size_t i0 = static_cast<size_t>(lb) * per, i1 = min(n16, i0 + per);
    const stream4_t* p = reinterpret_cast<const stream4_t*>(src);
    unsigned acc = 0;
    size_t i = i0 + threadIdx.x;
    for (; i + 3 * kBlock < i1; i += 4 * kBlock) {
        const stream4_t a = __builtin_nontemporal_load(p + i);
        const stream4_t b = __builtin_nontemporal_load(p + i + kBlock);
        const stream4_t c = __builtin_nontemporal_load(p + i + 2 * kBlock);
        const stream4_t d = __builtin_nontemporal_load(p + i + 3 * kBlock);
        const stream4_t t = a ^ b ^ c ^ d;
        acc ^= t.x ^ t.y ^ t.z ^ t.w;
    }
    for (; i < i1; i += kBlock) {
        const stream4_t t = __builtin_nontemporal_load(p + i);
        acc ^= t.x ^ t.y ^ t.z ^ t.w;
    }
    if (acc == 0x9E3779B9u) sink[lb] = acc;  // data-dependent, almost never taken
}

__global__ __launch_bounds__(kBlock, 8) void k_stream_copy(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                        size_t n16) {
    const int lb = logical_block();
    const size_t per = (n16 + gridDim.x - 1) / gridDim.x;
    const size_t i0 = static_cast<size_t>(lb) * per, i1 = min(n16, i0 + per);
    for (size_t i = i0 + threadIdx.x; i < i1; i += kBlock) dst[i] = src[i];
}

// The copy with non-temporal stores (NTL: and non-temporal loads), four
// 16-byte moves in flight per thread: the written lines stream out instead
// of waiting as dirty lines in the L2 for the kernel's end.
template <bool NTL>
__global__ __launch_bounds__(kBlock, 8) void k_stream_copy_nt(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                           size_t n16) {
    const int lb = logical_block();
    const size_t per = (n16 + gridDim.x - 1) / gridDim.x;
    const size_t i0 = static_cast<size_t>(lb) * per, i1 = min(n16, i0 + per);
    const stream4_t* p = reinterpret_cast<const stream4_t*>(src);
    stream4_t* q = reinterpret_cast<stream4_t*>(dst);
    size_t i = i0 + threadIdx.x;
    for (; i + 3 * kBlock < i1; i += 4 * kBlock) {
        stream4_t v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = NTL ? __builtin_nontemporal_load(p + i + k * kBlock) : p[i + k * kBlock];
#pragma unroll
        for (int k = 0; k < 4; ++k) __builtin_nontemporal_store(v[k], q + i + k * kBlock);
    }
    for (; i < i1; i += kBlock) {
        const stream4_t v = NTL ? __builtin_nontemporal_load(p + i) : p[i];
        __builtin_nontemporal_store(v, q + i);
    }
}

template <bool IDX16, int VMODE>
__global__ __launch_bounds__(kBlock, 8) void k_spmv_f64(Launch L, SellView A, const double* __restrict__ x,
                                                     double* __restrict__ y) {
    const WorkItem w = work_item(L, logical_block());
    XP_WAVE_LANE
    for (int sl = w.slice_begin + wave; sl < w.slice_end; sl += kWaves)
        y[sl * kC + lane] = sell_row_f64<IDX16, VMODE>(A, sl, lane, x);
}

// r = b - A x in FP64 and the partials of r^T r.
template <bool IDX16, int VMODE>
__global__ __launch_bounds__(kBlock, 8) void k_residual_f64(Launch L, SellView A, const double* __restrict__ x,
                                                         const double* __restrict__ b, double* __restrict__ r,
                                                         float* __restrict__ r32) {
    const int lb = logical_block();
    const WorkItem w = work_item(L, lb);
    if (w.sys < 0) return;
    XP_WAVE_LANE
    double v[1] = {0.0};
    for (int sl = w.slice_begin + wave; sl < w.slice_end; sl += kWaves) {
        const int row = sl * kC + lane;
        const double rv = b[row] - sell_row_f64<IDX16, VMODE>(A, sl, lane, x);
        r[row] = rv;
        if (r32 != nullptr) r32[row] = static_cast<float>(rv);
        v[0] += rv * rv;
    }
    block_sum<1>(v);
    if (threadIdx.x == 0) put_partial(L, kSlotAux, lb, v[0]);
}

// One workgroup per system: out[sys] = sum of the kSlotAux partials.
__global__ __launch_bounds__(kBlock, 8) void k_finish_aux(Launch L, double* out) {
    const int sys = static_cast<int>(blockIdx.x);
    double t[1];
    sum_partials<1>(L, sys, {kSlotAux}, t);
    if (threadIdx.x == 0) out[sys] = t[0];
}

// Residual replacement (reliable updates): the FP32 recursion's r and z are
// replaced by the true residual of the FP64 solution, r = fp32(b - A x64)
// with hi+lo (or exact) values, and z = M r; the rz / rr partials the next
// scalar step consumes are rewritten from them (as k_pcg_update_xrz writes
// them) and the Aux partials get the FP64 true r^T r.  p and ap are kept.
template <bool IDX16, int VMODE>
__global__ __launch_bounds__(kBlock, 8) void k_pcg_replace(Launch L, SellView A, const double* __restrict__ x,
                                                        const double* __restrict__ b,
                                                        const float* __restrict__ minv,
                                                        const BlockPre* __restrict__ blockM, float* __restrict__ r,
                                                        float* __restrict__ z) {
    const int lb = logical_block();
    const WorkItem w = work_item(L, lb);
    if (w.sys < 0) return;
    XP_WAVE_LANE
    double v[3] = {0.0, 0.0, 0.0};
    for (int sl = w.slice_begin + wave; sl < w.slice_end; sl += kWaves) {  // wave-uniform: block_apply shuffles
        const int row = sl * kC + lane;
        const double rv = b[row] - sell_row_f64<IDX16, VMODE>(A, sl, lane, x);
        const float r32 = static_cast<float>(rv);
        const float zv = blockM != nullptr ? block_apply(blockM, row, r32)
                                           : static_cast<float>(static_cast<double>(minv[row]) * static_cast<double>(r32));
        r[row] = r32;
        z[row] = zv;
        v[0] += static_cast<double>(r32) * static_cast<double>(zv);
        v[1] += static_cast<double>(r32) * static_cast<double>(r32);
        v[2] += rv * rv;
    }
    block_sum<3>(v);
    if (threadIdx.x == 0) {
        put_partial(L, kSlotRz, lb, v[0]);
        put_partial(L, kSlotRr, lb, v[1]);
        put_partial(L, kSlotAux, lb, v[2]);
    }
}

// One workgroup per system: out[3 sys + 0..2] = the replaced rz, rr and the
// FP64 true r^T r, summed in workgroup order.
__global__ __launch_bounds__(kBlock, 8) void k_replace_sums(Launch L, double* out) {
    const int sys = static_cast<int>(blockIdx.x);
    double t[3];
    sum_partials<3>(L, sys, {kSlotRz, kSlotRr, kSlotAux}, t);
    if (threadIdx.x != 0) return;
    out[3 * sys + 0] = t[0];
    out[3 * sys + 1] = t[1];
    out[3 * sys + 2] = t[2];
}

// x64 += (double) d; d = 0 (folds the FP32 correction into the FP64 solution).
__global__ void k_fold_f32_into_f64(int n, float* __restrict__ d, double* __restrict__ x) {
    const int i = static_cast<int>(blockIdx.x) * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
    if (i < n) {
        x[i] += static_cast<double>(d[i]);
        d[i] = 0.0f;
    }
}

__global__ void k_axpy_f64_f32(int n, const float* __restrict__ d, double* __restrict__ x) {
    const int i = static_cast<int>(blockIdx.x) * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
    if (i < n) x[i] += static_cast<double>(d[i]);
}

// ---------------------------------------------------------------------------
// Natural-order FP64 kernels behind the drop-in C-ABI.  Same formulas as
// archived/kernels/cg_kernels.cpp:73-158.  Each writes one partial per
// workgroup (partials[k * grid + block]); k_flat_finish then sums them in
// block order into out[0..NV-1].

__device__ __forceinline__ void flat_range(int n, int& i0, int& i1) {
    const int per = (n + static_cast<int>(gridDim.x) - 1) / static_cast<int>(gridDim.x);
    i0 = min(n, static_cast<int>(blockIdx.x) * per);
    i1 = min(n, i0 + per);
}

template <int NV>
__device__ __forceinline__ void flat_put(double* partials, double (&v)[NV]) {
    block_sum<NV>(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) partials[static_cast<size_t>(k) * gridDim.x + blockIdx.x] = v[k];
    }
}

template <int NV>
__global__ __launch_bounds__(kBlock, 8) void k_flat_finish(const double* partials, int nblocks, double* out) {
    double acc[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = 0.0;
    for (int b = static_cast<int>(threadIdx.x); b < nblocks; b += kBlock) {
#pragma unroll
        for (int k = 0; k < NV; ++k) acc[k] += partials[static_cast<size_t>(k) * nblocks + b];
    }
    block_sum<NV>(acc);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) out[k] = acc[k];
    }
}

__global__ __launch_bounds__(kBlock, 8) void k_init_pcg_f64(int n, const double* __restrict__ b,
                                                         const double* __restrict__ ax,
                                                         const double* __restrict__ m_inv,
                                                         double* __restrict__ r, double* __restrict__ z,
                                                         double* __restrict__ p, double* partials) {
    int i0, i1;
    flat_range(n, i0, i1);
    double v[2] = {0.0, 0.0};
    for (int i = i0 + static_cast<int>(threadIdx.x); i < i1; i += kBlock) {
        const double rv = b[i] - ax[i];
        const double zv = m_inv[i] * rv;
        r[i] = rv;
        z[i] = zv;
        p[i] = zv;
        v[0] += rv * zv;
        v[1] += rv * rv;
    }
    flat_put<2>(partials, v);
}

__global__ __launch_bounds__(kBlock, 8) void k_dot_f64(int n, const double* __restrict__ a,
                                                    const double* __restrict__ b, double* partials) {
    int i0, i1;
    flat_range(n, i0, i1);
    double v[1] = {0.0};
    for (int i = i0 + static_cast<int>(threadIdx.x); i < i1; i += kBlock) v[0] += a[i] * b[i];
    flat_put<1>(partials, v);
}

__global__ __launch_bounds__(kBlock, 8) void k_update_xrz_f64(int n, double* __restrict__ x,
                                                           const double* __restrict__ p,
                                                           double* __restrict__ r,
                                                           const double* __restrict__ ap,
                                                           const double* __restrict__ m_inv,
                                                           double* __restrict__ z, double alpha,
                                                           double* partials) {
    int i0, i1;
    flat_range(n, i0, i1);
    double v[2] = {0.0, 0.0};
    for (int i = i0 + static_cast<int>(threadIdx.x); i < i1; i += kBlock) {
        x[i] += alpha * p[i];
        const double rv = r[i] - alpha * ap[i];
        const double zv = m_inv[i] * rv;
        r[i] = rv;
        z[i] = zv;
        v[0] += rv * zv;
        v[1] += rv * rv;
    }
    flat_put<2>(partials, v);
}

__global__ void k_update_p_f64(int n, const double* __restrict__ z, double* __restrict__ p, double beta) {
    const int i = static_cast<int>(blockIdx.x) * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
    if (i < n) p[i] = z[i] + beta * p[i];
}

template <typename S, typename D>
__global__ void k_gather(int n_pad, const int* __restrict__ perm, const S* __restrict__ src,
                         D* __restrict__ dst) {
    const int i = static_cast<int>(blockIdx.x) * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
    if (i >= n_pad) return;
    const int g = perm[i];
    dst[i] = g >= 0 ? static_cast<D>(src[g]) : static_cast<D>(0);
}

template <typename S, typename D>
__global__ void k_scatter(int n_pad, const int* __restrict__ perm, const S* __restrict__ src,
                          D* __restrict__ dst) {
    const int i = static_cast<int>(blockIdx.x) * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
    if (i >= n_pad) return;
    const int g = perm[i];
    if (g >= 0) dst[g] = static_cast<D>(src[i]);
}

__global__ void k_convert_f2d(int n, const float* __restrict__ src, double* __restrict__ dst) {
    const int i = static_cast<int>(blockIdx.x) * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
    if (i < n) dst[i] = static_cast<double>(src[i]);
}

template <typename T>
__global__ void k_fill(size_t n, T v, T* __restrict__ dst) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = v;
}

int blocks_for(size_t n, int bs) { return static_cast<int>((n + static_cast<size_t>(bs) - 1) / static_cast<size_t>(bs)); }


// ---------------------------------------------------------------------------
// Jacobi iteration x_new = D^-1 (b - R x), A = D + R
// (DLC/Cuper-jacobi-iteration/host/main.cpp:957-1053, docs/jacobi_iteration.md:1-60).
// FP32 throughout, as the reference's VALUE_TYPE float.  R x is summed per
// row in CSR order from 0, one rounding per product and one per sum (no
// fused multiply-add: contraction is switched off in the helpers below),
// skipping every entry on the diagonal: the SELL slots
// after the first diagonal entry keep the CSR order, so each row's R x, and
// with it x_new, is the reference's bit for bit.

// NaN-propagating maximum: a NaN in b wins, a NaN in a is kept.
__device__ __forceinline__ double nan_max(double a, double b) { return (b > a || b != b) ? b : a; }

// Maximum of partial slot `slot` over a system's logical blocks; every
// thread receives it; ends with a barrier.
__device__ __forceinline__ double max_partials(const Launch& L, int sys, int slot) {
    __shared__ double red[kWaves];
    const int2 range = sys_range(L, sys);
    double m = 0.0;
    for (int b = range.x + static_cast<int>(threadIdx.x); b < range.y; b += kBlock)
        m = nan_max(m, L.partials[static_cast<size_t>(slot) * L.grid + b]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = nan_max(m, __shfl_xor(m, off, 64));
    const int wave = static_cast<int>(threadIdx.x) >> 6;
    if ((threadIdx.x & 63) == 0) red[wave] = m;
    __syncthreads();
    m = red[0];
#pragma unroll
    for (int k = 1; k < kWaves; ++k) m = nan_max(m, red[k]);
    __syncthreads();
    return m;
}

// acc + v * x with two roundings.  The HIP default contracts a product and
// a sum into one fused multiply-add (a single rounding); the pragma keeps
// them apart, as the reference's host loop computes them.
__device__ __forceinline__ float add_product_rn(float acc, float v, float x) {
#pragma clang fp contract(off)
    const float p = v * x;
    return acc + p;
}

// next = (b - rx) * dinv and |next - x_old|, each operation rounded (no
// product fused into the following difference).
__device__ __forceinline__ float jacobi_next(float b, float rx, float dinv, float xold, float& diff) {
#pragma clang fp contract(off)
    const float next = (b - rx) * dinv;
    diff = fabsf(next - xold);
    return next;
}

// R x of one SELL row: the off-diagonal slots in CSR order, entries whose
// column is the row itself (padding, duplicate diagonal entries) skipped.
template <bool IDX16>
__device__ __forceinline__ float sell_row_offdiag_seq(const SellView& A, int slice, int lane,
                                                      const float* __restrict__ x) {
    const int j0 = A.slice_ptr[slice];
    const int w = A.slice_ptr[slice + 1] - j0;
    const int row = slice * kC + lane;
    const size_t e0 = static_cast<size_t>(j0) * kC + static_cast<size_t>(lane);
    const int base = A.slice_base[slice];
    float acc = 0.0f;
    if (w <= kFastW) {
        // All loads first.  Slots past the width re-read the last real slot
        // (slot 0, whose values are not otherwise needed, only when the
        // slice has no off-diagonal column) and are masked below.
        float v[kFastW], xv[kFastW];
        int c[kFastW];
        const int last = w > 1 ? w - 1 : 0;
#pragma unroll
        for (int k = 1; k < kFastW; ++k) {
            const size_t e = e0 + static_cast<size_t>(k < w ? k : last) * kC;
            v[k] = A.val[e];
            c[k] = k < w ? (IDX16 ? base + static_cast<int>(A.col16[e]) : A.col32[e]) : row;
        }
#pragma unroll
        for (int k = 1; k < kFastW; ++k) xv[k] = x[c[k]];
#pragma unroll
        for (int k = 1; k < kFastW; ++k)
            acc = (k < w && c[k] != row) ? add_product_rn(acc, v[k], xv[k]) : acc;
        return acc;
    }
    size_t e = e0;
    for (int j = 1; j < w; ++j) {
        e += kC;
        const int c = IDX16 ? base + static_cast<int>(A.col16[e]) : A.col32[e];
        if (c != row) acc = add_product_rn(acc, A.val[e], x[c]);
    }
    return acc;
}

// R x for U slices of one wavefront at once (its slices sl_first,
// sl_first + kWaves, ...), every slice at most FW columns wide: all value,
// column and x loads of the U slices are issued before the first sum.
// Slices past sl_last re-read sl_last (not written by the caller); slots
// past a slice's width re-read its last slot (slot 0 when it has no
// off-diagonal column) and are masked, as are entries on the diagonal.
template <bool IDX16, int FW, int U>
__device__ __forceinline__ void jacobi_group(const SellView& A, int sl_first, int sl_last, int lane,
                                             const float* __restrict__ x, const float* __restrict__ b,
                                             const float* __restrict__ dinv, float (&rx)[U], float (&bv)[U],
                                             float (&dv)[U], float (&xr)[U]) {
    int row[U], w[U], base[U], last[U];
    size_t e0[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int s = min(sl_first + u * kWaves, sl_last);
        const int j0 = A.slice_ptr[s];
        w[u] = A.slice_ptr[s + 1] - j0;
        row[u] = s * kC + lane;
        e0[u] = static_cast<size_t>(j0) * kC + static_cast<size_t>(lane);
        base[u] = A.slice_base[s];
        last[u] = w[u] > 1 ? w[u] - 1 : 0;
    }
    float v[U][FW];
    int c[U][FW];
#pragma unroll
    for (int u = 0; u < U; ++u) {
#pragma unroll
        for (int k = 1; k < FW; ++k) {
            const size_t e = e0[u] + static_cast<size_t>(k < w[u] ? k : last[u]) * kC;
            v[u][k] = A.val[e];
            c[u][k] = k < w[u] ? (IDX16 ? base[u] + static_cast<int>(A.col16[e]) : A.col32[e]) : row[u];
        }
        bv[u] = b[row[u]];
        dv[u] = dinv[row[u]];
        xr[u] = x[row[u]];
    }
    float xg[U][FW];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
        for (int k = 1; k < FW; ++k) xg[u][k] = x[c[u][k]];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        float acc = 0.0f;
#pragma unroll
        for (int k = 1; k < FW; ++k)
            acc = (k < w[u] && c[u][k] != row[u]) ? add_product_rn(acc, v[u][k], xg[u][k]) : acc;
        rx[u] = acc;
    }
}

// One Jacobi update.  Reads state slot Sc, writes Sn.  First the stop test
// of the previous update (its max |x_new - x_old| from the partials):
// non-finite -> breakdown; use_tau and diff <= tau -> converged; then
// iters == max_iters -> max_iter.  Update j runs in launch j and reads x
// from buffer j % 2 (xo), writing buffer (j + 1) % 2 (xn): the caller passes
// them by launch parity, so each wavefront loads and sums its first U
// slices before the stop test is known (only the writes wait for it).  A
// stopped system's kernels are no-ops.  FW > 0: every slice is at most FW
// columns wide and U slices go through jacobi_group together; FW = 0: any
// width, one slice at a time.
template <bool IDX16, int FW, int U>
__global__ __launch_bounds__(kBlock, 8) void k_jacobi_iter(Launch L, SellView A, const JacState* Sc, JacState* Sn,
                                                         const float* __restrict__ b,
                                                         const float* __restrict__ dinv,
                                                         const float* __restrict__ xo, float* __restrict__ xn,
                                                         int slot_read, int slot_write) {
    __shared__ JacState s_sh;
    const int lb = logical_block();
    const WorkItem w = work_item(L, lb);
    if (w.sys < 0) return;
    const bool leader_block = lb == sys_range(L, w.sys).x;
    if (threadIdx.x == 0) s_sh = Sc[w.sys];
    XP_WAVE_LANE
    const int sl0 = w.slice_begin + wave;
    // Work items are never empty; the guard keeps an empty one from
    // reading outside the matrix.
    const bool any = w.slice_end > w.slice_begin;
    const int sl_last = w.slice_end - 1;
    float rx[U] = {}, bv[U] = {}, dv[U] = {}, xr[U] = {};
    if (any) {
        if (FW > 0) {
            jacobi_group<IDX16, (FW > 0 ? FW : 1), U>(A, sl0, sl_last, lane, xo, b, dinv, rx, bv, dv, xr);
        } else {
            const int s = min(sl0, sl_last);
            const int row = s * kC + lane;
            rx[0] = sell_row_offdiag_seq<IDX16>(A, s, lane, xo);
            bv[0] = b[row];
            dv[0] = dinv[row];
            xr[0] = xo[row];
        }
    }
    const double prev = max_partials(L, w.sys, slot_read);  // ends with a barrier
    if (!s_sh.active) {
        if (leader_block && threadIdx.x == 0) Sn[w.sys] = s_sh;
        return;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        JacState& s = s_sh;
        if (s.pending) {
            s.pending = 0;
            s.diff = prev;
            if (!isfinite(prev)) {
                s.status = kStatusBreakdown;
                s.active = 0;
            } else if (s.use_tau && prev <= s.tau) {
                s.status = kStatusConverged;
                s.active = 0;
            }
        }
        if (s.active && s.iters >= s.max_iters) {
            s.status = kStatusMaxIter;
            s.active = 0;
        }
        if (s.active) {
            s.iters += 1;
            s.pending = 1;
        }
        if (leader_block) Sn[w.sys] = s;
    }
    __syncthreads();
    if (!s_sh.active) return;
    constexpr int kStep = (FW > 0 ? U : 1) * kWaves;
    double m = 0.0;
    for (int sl = sl0; sl < w.slice_end; sl += kStep) {
        if (sl != sl0) {   // the first group was loaded ahead of the stop test
            if (FW > 0) {
                jacobi_group<IDX16, (FW > 0 ? FW : 1), U>(A, sl, sl_last, lane, xo, b, dinv, rx, bv, dv, xr);
            } else {
                const int row = sl * kC + lane;
                rx[0] = sell_row_offdiag_seq<IDX16>(A, sl, lane, xo);
                bv[0] = b[row];
                dv[0] = dinv[row];
                xr[0] = xo[row];
            }
        }
#pragma unroll
        for (int u = 0; u < (FW > 0 ? U : 1); ++u) {
            const int su = sl + u * kWaves;
            if (su < w.slice_end) {
                float diff;
                xn[su * kC + lane] = jacobi_next(bv[u], rx[u], dv[u], xr[u], diff);
                m = nan_max(m, static_cast<double>(diff));
            }
        }
    }
    __shared__ double red[kWaves];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = nan_max(m, __shfl_xor(m, off, 64));
    if (lane == 0) red[wave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = red[0];
        for (int k = 1; k < kWaves; ++k) t = nan_max(t, red[k]);
        put_partial(L, slot_write, lb, t);
    }
}


// One Jacobi update of the fixed-count contract (host/main.cpp:1011-1053:
// max_iters updates, no stop test): no state, no wait on other
// workgroups; each workgroup leaves its max |x_new - x_old| in partial
// slot 0, which k_jacobi_finish reads after the last update.  Buffers by
// launch parity, as k_jacobi_iter.
template <bool IDX16, int FW, int U>
__global__ __launch_bounds__(kBlock, 8) void k_jacobi_sweep(Launch L, SellView A, const float* __restrict__ b,
                                                          const float* __restrict__ dinv,
                                                          const float* __restrict__ xo, float* __restrict__ xn) {
    const int lb = logical_block();
    const WorkItem w = work_item(L, lb);
    if (w.sys < 0) return;
    XP_WAVE_LANE
    const int sl_last = w.slice_end - 1;
    constexpr int kStep = (FW > 0 ? U : 1) * kWaves;
    double m = 0.0;
    for (int sl = w.slice_begin + wave; sl < w.slice_end; sl += kStep) {
        float rx[U] = {}, bv[U] = {}, dv[U] = {}, xr[U] = {};
        if (FW > 0) {
            jacobi_group<IDX16, (FW > 0 ? FW : 1), U>(A, sl, sl_last, lane, xo, b, dinv, rx, bv, dv, xr);
        } else {
            const int row = sl * kC + lane;
            rx[0] = sell_row_offdiag_seq<IDX16>(A, sl, lane, xo);
            bv[0] = b[row];
            dv[0] = dinv[row];
            xr[0] = xo[row];
        }
#pragma unroll
        for (int u = 0; u < (FW > 0 ? U : 1); ++u) {
            const int su = sl + u * kWaves;
            if (su < w.slice_end) {
                float diff;
                xn[su * kC + lane] = jacobi_next(bv[u], rx[u], dv[u], xr[u], diff);
                m = nan_max(m, static_cast<double>(diff));
            }
        }
    }
    __shared__ double red[kWaves];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = nan_max(m, __shfl_xor(m, off, 64));
    if (lane == 0) red[wave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = red[0];
        for (int k = 1; k < kWaves; ++k) t = nan_max(t, red[k]);
        put_partial(L, 0, lb, t);
    }
}

// After `iters` fixed-count updates: one workgroup per system turns the
// last update's partial maxima into the system's state (diff; max_iter, or
// breakdown when that difference is not finite).  iters = 0: no update ran
// and diff is 0.
__global__ __launch_bounds__(kBlock, 1) void k_jacobi_finish(Launch L, JacState* S, int iters) {
    const int sys = static_cast<int>(blockIdx.x);
    const double m = iters > 0 ? max_partials(L, sys, 0) : 0.0;
    if (threadIdx.x == 0) {
        JacState s = S[sys];
        s.diff = m;
        s.iters = iters;
        s.pending = 0;
        s.active = 0;
        s.status = isfinite(m) ? kStatusMaxIter : kStatusBreakdown;
        S[sys] = s;
    }
}

// ---------------------------------------------------------------------------
// Stream-ordered FP32 device operations: the five fused operations of
// archived/kernels/cg_kernels.cpp:73-158 on device vectors in solve order,
// for callers that drive their own iteration.  Each value is computed in
// FP64 from the FP32 operands (no fused multiply-add) and rounded once to
// FP32; the dot products are FP64 sums of the exact FP32 products, one per
// workgroup into the partial slots, finished by k_op_finish in a fixed
// order.

__global__ __launch_bounds__(kBlock, 8) void k_op_init(Launch L, const float* __restrict__ b,
                                                     const float* __restrict__ ax,
                                                     const float* __restrict__ minv, float* __restrict__ r,
                                                     float* __restrict__ z, float* __restrict__ p) {
#pragma clang fp contract(off)
    const int lb = logical_block();
    const WorkItem w = work_item(L, lb);
    if (w.sys < 0) return;
    double v[2] = {0.0, 0.0};
    for (int row = w.slice_begin * kC + static_cast<int>(threadIdx.x); row < w.slice_end * kC; row += kBlock) {
        const float rv = static_cast<float>(static_cast<double>(b[row]) - static_cast<double>(ax[row]));
        const float zv = static_cast<float>(static_cast<double>(minv[row]) * static_cast<double>(rv));
        r[row] = rv;
        z[row] = zv;
        p[row] = zv;
        v[0] += static_cast<double>(rv) * static_cast<double>(zv);
        v[1] += static_cast<double>(rv) * static_cast<double>(rv);
    }
    block_sum<2>(v);
    if (threadIdx.x == 0) {
        put_partial(L, kSlotRz, lb, v[0]);
        put_partial(L, kSlotRr, lb, v[1]);
    }
}

__global__ __launch_bounds__(kBlock, 8) void k_op_dot(Launch L, const float* __restrict__ a,
                                                    const float* __restrict__ b) {
    const int lb = logical_block();
    const WorkItem w = work_item(L, lb);
    if (w.sys < 0) return;
    double v[1] = {0.0};
    for (int row = w.slice_begin * kC + static_cast<int>(threadIdx.x); row < w.slice_end * kC; row += kBlock)
        v[0] += static_cast<double>(a[row]) * static_cast<double>(b[row]);
    block_sum<1>(v);
    if (threadIdx.x == 0) put_partial(L, kSlotAux, lb, v[0]);
}

__global__ __launch_bounds__(kBlock, 8) void k_op_update_xrz(Launch L, float* __restrict__ x,
                                                           const float* __restrict__ p, float* __restrict__ r,
                                                           const float* __restrict__ ap,
                                                           const float* __restrict__ minv, float* __restrict__ z,
                                                           double alpha) {
#pragma clang fp contract(off)
    const int lb = logical_block();
    const WorkItem w = work_item(L, lb);
    if (w.sys < 0) return;
    double v[2] = {0.0, 0.0};
    for (int row = w.slice_begin * kC + static_cast<int>(threadIdx.x); row < w.slice_end * kC; row += kBlock) {
        const double ap_a = alpha * static_cast<double>(ap[row]);
        const double p_a = alpha * static_cast<double>(p[row]);
        const float rv = static_cast<float>(static_cast<double>(r[row]) - ap_a);
        const float zv = static_cast<float>(static_cast<double>(minv[row]) * static_cast<double>(rv));
        x[row] = static_cast<float>(static_cast<double>(x[row]) + p_a);
        r[row] = rv;
        z[row] = zv;
        v[0] += static_cast<double>(rv) * static_cast<double>(zv);
        v[1] += static_cast<double>(rv) * static_cast<double>(rv);
    }
    block_sum<2>(v);
    if (threadIdx.x == 0) {
        put_partial(L, kSlotRz, lb, v[0]);
        put_partial(L, kSlotRr, lb, v[1]);
    }
}

__global__ __launch_bounds__(kBlock, 8) void k_op_update_p(Launch L, const float* __restrict__ z,
                                                         float* __restrict__ p, double beta) {
#pragma clang fp contract(off)
    const WorkItem w = work_item(L, logical_block());
    if (w.sys < 0) return;
    for (int row = w.slice_begin * kC + static_cast<int>(threadIdx.x); row < w.slice_end * kC; row += kBlock) {
        const double bp = beta * static_cast<double>(p[row]);
        p[row] = static_cast<float>(static_cast<double>(z[row]) + bp);
    }
}

// One workgroup per system: out[sys * nslot + k] = the sum of partial slot
// slot0 + k over the system's workgroups.
__global__ __launch_bounds__(kBlock, 1) void k_op_finish(Launch L, int slot0, int nslot, double* __restrict__ out) {
    const int sys = static_cast<int>(blockIdx.x);
    if (nslot == 1) {
        double t[1];
        sum_partials<1>(L, sys, {slot0}, t);
        if (threadIdx.x == 0) out[sys] = t[0];
    } else {
        double t[2];
        sum_partials<2>(L, sys, {slot0, slot0 + 1}, t);
        if (threadIdx.x == 0) {
            out[2 * sys] = t[0];
            out[2 * sys + 1] = t[1];
        }
    }
}

// ---------------------------------------------------------------------------
// Multilevel (aggregation) preconditioner.  Every level is a SELL matrix
// with its own 8-row smoother blocks; the coarse space of level l is a set
// of aggregates of its rows (plain aggregation, below), and row u of level
// l + 1 is aggregate u.  P is piecewise constant on the aggregates, the
// coarse operator is P^T A P, and one V-cycle is
//   z = omega M r;  r' = P^T (r - A z);  z += gamma P V(r');  z += omega M (r - A z)
// with M the level's block factors and the coarsest level solved by a dense
// inverse.  The cycle is symmetric (the same smoother before and after).
// Each coarse row carries the list of its aggregate's rows as positions in
// the finer level's solve order (mp / mpos): the restriction is a gather
// over that list, the prolongation a scatter over it.  Vectors, values and
// sums are FP32 (a preconditioner needs no more); every kernel returns at
// once when the system has stopped.

// w = r - c (A z) for every solve row of the level (padding rows: r and z
// are 0 there, so w = 0).
template <bool IDX16, int V>
__global__ __launch_bounds__(kBlock, 8) void k_mg_restrict(Launch L, SellView A, const SysState* st,
                                                        const float* __restrict__ r, const float* __restrict__ z,
                                                        float c, float* __restrict__ wout) {
    if (st != nullptr && !st[0].active) return;
    const WorkItem w = work_item(L, logical_block());
    if (w.sys < 0) return;
    XP_WAVE_LANE
    for (int sl = w.slice_begin + wave; sl < w.slice_end; sl += kWaves) {
        const int row = sl * kC + lane;
        float zd;
        const float az = sell_row_f32<IDX16, V>(A, sl, lane, z, zd);
        wout[row] = fmaf(-c, az, r[row]);
    }
}

// The coarse levels' small matrices give a workgroup at most one slice: its
// 16 wavefronts then share the slice's slots (wavefront k takes slots k,
// k + 16, ...), and wavefront 0 adds the 16 partial sums in wavefront order
// (LDS) and finishes the rows.  Returns the row sum in wavefront 0 (0
// elsewhere) and xdiag = x[row] there.
template <bool IDX16>
__device__ __forceinline__ float sell_row_split(const SellView& A, int slice, int wave, int lane,
                                                const float* __restrict__ x, float& xdiag, float (*part)[kC]) {
    const int j0 = A.slice_ptr[slice];
    const int wdt = A.slice_ptr[slice + 1] - j0;
    const int row = slice * kC + lane;
    const int base = A.slice_base[slice];
    float acc = 0.0f;
    for (int j = wave; j < wdt; j += kWaves) {
        const size_t e = (static_cast<size_t>(j0) + static_cast<size_t>(j)) * kC + static_cast<size_t>(lane);
        const int c = j == 0 ? row : (IDX16 ? base + static_cast<int>(A.col16[e]) : A.col32[e]);
        acc = fmaf(A.val[e], x[c], acc);
    }
    part[wave][lane] = acc;
    __syncthreads();
    float sum = 0.0f;
    if (wave == 0) {
        for (int k = 0; k < kWaves; ++k) sum += part[k][lane];
        xdiag = x[row];
    }
    __syncthreads();  // part is reused by the next slice
    return sum;
}

template <bool IDX16>
__global__ __launch_bounds__(kBlock, 8) void k_mg_restrict_split(Launch L, SellView A, const SysState* st,
                                                              const float* __restrict__ r,
                                                              const float* __restrict__ z, float c,
                                                              float* __restrict__ wout) {
    __shared__ float part[kWaves][kC];
    if (st != nullptr && !st[0].active) return;
    const WorkItem w = work_item(L, logical_block());
    if (w.sys < 0) return;
    XP_WAVE_LANE
    for (int sl = w.slice_begin; sl < w.slice_end; ++sl) {
        float zd = 0.0f;
        const float az = sell_row_split<IDX16>(A, sl, wave, lane, z, zd, part);
        if (wave == 0) {
            const int row = sl * kC + lane;
            wout[row] = fmaf(-c, az, r[row]);
        }
    }
}

// z = omega M r (the smoother from a zero start).
__global__ __launch_bounds__(kBlock, 8) void k_mg_smooth(Launch L, const SysState* st,
                                                      const BlockPre* __restrict__ M, const float* __restrict__ r,
                                                      float omega, float* __restrict__ z) {
    if (st != nullptr && !st[0].active) return;
    const WorkItem w = work_item(L, logical_block());
    if (w.sys < 0) return;
    XP_WAVE_LANE
    for (int sl = w.slice_begin + wave; sl < w.slice_end; sl += kWaves) {
        const int row = sl * kC + lane;
        z[row] = omega * block_apply(M, row, r[row]);
    }
}

// The restriction's sum for coarse row `row`: the finer level's w over the
// row's aggregate, in list order (padding rows have empty lists: 0).  Eight
// members at a time: their positions, then their values, are loaded before
// the adds (the adds keep list order).
__device__ __forceinline__ float gather_members(const int* __restrict__ mp, const int* __restrict__ mpos,
                                                const float* __restrict__ wf, int row) {
    float s = 0.0f;
    const int e = mp[row + 1];
    for (int k = mp[row]; k < e; k += 8) {
        int idx[8];
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) idx[u] = k + u < e ? mpos[k + u] : -1;
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = idx[u] >= 0 ? wf[idx[u]] : 0.0f;
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (k + u < e) s += v[u];
    }
    return s;
}

// Coarse level: r = P^T w (the gather) and z = omega M r.
__global__ __launch_bounds__(kBlock, 8) void k_mg_smooth_gather(Launch L, const SysState* st,
                                                             const BlockPre* __restrict__ M,
                                                             const int* __restrict__ mp,
                                                             const int* __restrict__ mpos,
                                                             const float* __restrict__ wf, float* __restrict__ r,
                                                             float omega, float* __restrict__ z) {
    if (st != nullptr && !st[0].active) return;
    const WorkItem w = work_item(L, logical_block());
    if (w.sys < 0) return;
    XP_WAVE_LANE
    for (int sl = w.slice_begin + wave; sl < w.slice_end; sl += kWaves) {
        const int row = sl * kC + lane;
        const float rv = gather_members(mp, mpos, wf, row);
        r[row] = rv;
        z[row] = omega * block_apply(M, row, rv);
    }
}

// Prolongation into the finer level, done by the kernel that produces the
// coarse value e of row `row`: every row of its aggregate gets
// t = cz z + gamma e.
__device__ __forceinline__ void prolong_members(const MgUp& pu, int row, float gamma, float e) {
    const int end = pu.mp[row + 1];
    for (int k = pu.mp[row]; k < end; k += 8) {  // eight members' loads in flight at a time
        int idx[8];
        float zv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) idx[u] = k + u < end ? pu.mpos[k + u] : -1;
#pragma unroll
        for (int u = 0; u < 8; ++u) zv[u] = idx[u] >= 0 ? pu.z[idx[u]] : 0.0f;
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (idx[u] >= 0) pu.t[idx[u]] = fmaf(gamma, e, pu.cz * zv[u]);
    }
}

// zout = t + omega M (r - A t); RZ: the r^T zout partials of the PCG step
// (the solver's slot kSlotRz: the cycle's z replaces the update kernel's).
template <bool IDX16, int V, bool RZ>
__device__ __forceinline__ void mg_post(const Launch& L, const SellView& A, const SysState* st,
                                                    const BlockPre* __restrict__ M, const float* __restrict__ r,
                                                    const float* __restrict__ t, float omega,
                                                    float* __restrict__ zout, const MgUp& pu, float gamma) {
    if (st != nullptr && !st[0].active) return;
    const int lb = logical_block();
    const WorkItem w = work_item(L, lb);
    if (w.sys < 0) return;
    XP_WAVE_LANE
    double v[1] = {0.0};
    for (int sl = w.slice_begin + wave; sl < w.slice_end; sl += kWaves) {
        const int row = sl * kC + lane;
        float td;
        const float at = sell_row_f32<IDX16, V>(A, sl, lane, t, td);
        const float rv = r[row];
        const float zz = fmaf(omega, block_apply(M, row, rv - at), td);
        zout[row] = zz;
        if (RZ) v[0] += static_cast<double>(rv) * static_cast<double>(zz);
        if (pu.mp != nullptr) prolong_members(pu, row, gamma, zz);
    }
    if (RZ) {
        block_sum<1>(v);
        if (threadIdx.x == 0) put_partial(L, kSlotRz, lb, v[0]);
    }
}

// k_mg_post for one slice per workgroup (sell_row_split), no RZ partials.
template <bool IDX16>
__global__ __launch_bounds__(kBlock, 8) void k_mg_post_split(Launch L, SellView A, const SysState* st,
                                                          const BlockPre* __restrict__ M, const float* __restrict__ r,
                                                          const float* __restrict__ t, float omega,
                                                          float* __restrict__ zout, MgUp pu, float gamma) {
    __shared__ float part[kWaves][kC];
    if (st != nullptr && !st[0].active) return;
    const WorkItem w = work_item(L, logical_block());
    if (w.sys < 0) return;
    XP_WAVE_LANE
    for (int sl = w.slice_begin; sl < w.slice_end; ++sl) {
        float td = 0.0f;
        const float at = sell_row_split<IDX16>(A, sl, wave, lane, t, td, part);
        if (wave == 0) {
            const int row = sl * kC + lane;
            const float rv = r[row];
            const float zz = fmaf(omega, block_apply(M, row, rv - at), td);
            zout[row] = zz;
            if (pu.mp != nullptr) prolong_members(pu, row, gamma, zz);
        }
    }
}

// The dense coarsest level: every workgroup gathers the whole restricted
// residual rc = P^T w into LDS (nc <= kMaxCoarse), then e = Cinv rc with one
// wavefront per row, then the prolongation of e into the finer level (pu:
// its lanes scatter over the row's aggregate).
constexpr int kMaxCoarse = 4096;
__global__ __launch_bounds__(kBlock, 8) void k_mg_coarse(const SysState* st, int nc, const float* __restrict__ cinv,
                                                      const float* __restrict__ wf, float* __restrict__ ec,
                                                      MgUp pu, float gamma) {
    __shared__ float rc[kMaxCoarse];
    if (st != nullptr && !st[0].active) return;
    for (int j = static_cast<int>(threadIdx.x); j < nc; j += kBlock) rc[j] = gather_members(pu.mp, pu.mpos, wf, j);
    __syncthreads();
    const int row = static_cast<int>(blockIdx.x) * kWaves + __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) >> 6);
    const int lane = static_cast<int>(threadIdx.x) & 63;
    if (row >= nc) return;  // wave-uniform
    double s = 0.0;
    for (int j = lane; j < nc; j += kC) s += static_cast<double>(cinv[static_cast<size_t>(row) * nc + j]) * rc[j];
    s = wave_sum(s);  // every lane holds the sum
    const float e = static_cast<float>(s);
    if (lane == 0) ec[row] = e;
    const int end = pu.mp[row + 1];
    for (int k = pu.mp[row] + lane; k < end; k += kC) {
        const int i = pu.mpos[k];
        pu.t[i] = fmaf(gamma, e, pu.cz * pu.z[i]);
    }
}

template <bool IDX16, int V>
__global__ __launch_bounds__(kBlock, 8) void k_mg_post_norz(Launch L, SellView A, const SysState* st,
                                                         const BlockPre* __restrict__ M, const float* __restrict__ r,
                                                         const float* __restrict__ t, float omega,
                                                         float* __restrict__ zout, MgUp pu, float gamma) {
    mg_post<IDX16, V, false>(L, A, st, M, r, t, omega, zout, pu, gamma);
}
template <bool IDX16, int V>
__global__ __launch_bounds__(kBlock, 8) void k_mg_post_rz(Launch L, SellView A, const SysState* st,
                                                       const BlockPre* __restrict__ M, const float* __restrict__ r,
                                                       const float* __restrict__ t, float omega,
                                                       float* __restrict__ zout) {
    mg_post<IDX16, V, true>(L, A, st, M, r, t, omega, zout, MgUp{}, 0.0f);
}

// ---- set-up of the hierarchy: plain aggregation ----
// Strength: j != i is a strong neighbour of i when a_ij^2 >= theta^2 |a_ii|
// |a_jj| (FP64; a_ii is the sum of the row's diagonal entries in CSR order).
// Roots are a maximal distance-2 independent set of the strength graph,
// chosen in Luby rounds: an undecided row becomes a root when its key is
// the largest among the undecided rows within strong distance 2, then the
// rows within distance 2 of a root are excluded.  Root k (index order) is
// aggregate k, its strong neighbours join it, every other row joins the
// aggregate of its strongest joined neighbour (largest |a_ij|, then the
// smallest j), rows left over become aggregates of their own in index
// order.  Every decision of a round depends on the previous round's state
// only, so the result does not depend on the schedule.

// The key of row i: the high 32 bits of MurmurHash3's 64-bit finaliser of
// i + salt * 0x9E3779B97F4A7C15, above i + 1 (unique and nonzero).
__device__ __forceinline__ unsigned long long agg_key(int i, int salt) {
    unsigned long long x = static_cast<unsigned long long>(i) +
                           static_cast<unsigned long long>(salt) * 0x9E3779B97F4A7C15ull;
    x ^= x >> 33;
    x *= 0xFF51AFD7ED558CCDull;
    x ^= x >> 33;
    x *= 0xC4CEB9FE1A85EC53ull;
    x ^= x >> 33;
    return ((x >> 32) << 32) | static_cast<unsigned long long>(i + 1);
}

__global__ void k_row_absdiag(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                              const double* __restrict__ va, double* __restrict__ d) {
    const int i = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (i >= n) return;
    double s = 0.0;
    for (int e = rp[i]; e < rp[i + 1]; ++e)
        if (ci[e] == i) s += va[e];
    d[i] = fabs(s);
}

__global__ void k_agg_strength(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                               const double* __restrict__ va, const double* __restrict__ d, double th2,
                               unsigned char* __restrict__ strong) {
#pragma clang fp contract(off)
    const int i = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (i >= n) return;
    const double di = d[i];
    for (int e = rp[i]; e < rp[i + 1]; ++e) {
        const int j = ci[e];
        const double v = va[e];
        strong[e] = (j != i && v * v >= th2 * (di * d[j])) ? 1 : 0;
    }
}

// state: 0 undecided, 1 root, 2 excluded.
__global__ void k_mis_max1(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                           const unsigned char* __restrict__ strong, const unsigned char* __restrict__ state,
                           int salt, unsigned long long* __restrict__ m1) {
    const int i = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (i >= n) return;
    unsigned long long best = state[i] == 0 ? agg_key(i, salt) : 0ull;
    for (int e = rp[i]; e < rp[i + 1]; ++e) {
        const int j = ci[e];
        if (strong[e] && state[j] == 0) {
            const unsigned long long k = agg_key(j, salt);
            best = k > best ? k : best;
        }
    }
    m1[i] = best;
}

__global__ void k_mis_root(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                           const unsigned char* __restrict__ strong, const unsigned long long* __restrict__ m1,
                           int salt, unsigned char* __restrict__ state) {
    const int i = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (i >= n || state[i] != 0) return;
    unsigned long long best = m1[i];
    for (int e = rp[i]; e < rp[i + 1]; ++e)
        if (strong[e]) best = m1[ci[e]] > best ? m1[ci[e]] : best;
    if (best == agg_key(i, salt)) state[i] = 1;
}

__global__ void k_mis_near(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                           const unsigned char* __restrict__ strong, const unsigned char* __restrict__ state,
                           unsigned char* __restrict__ near) {
    const int i = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (i >= n) return;
    int f = state[i] == 1;
    for (int e = rp[i]; e < rp[i + 1] && !f; ++e)
        if (strong[e] && state[ci[e]] == 1) f = 1;
    near[i] = static_cast<unsigned char>(f);
}

// Excludes the undecided rows near a root; *left += the rows still undecided.
__global__ __launch_bounds__(256) void k_mis_exclude(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                     const unsigned char* __restrict__ strong,
                                                     const unsigned char* __restrict__ near,
                                                     unsigned char* __restrict__ state, int* __restrict__ left) {
    const int i = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    int und = 0;
    if (i < n && state[i] == 0) {
        int f = near[i];
        for (int e = rp[i]; e < rp[i + 1] && !f; ++e)
            if (strong[e] && near[ci[e]]) f = 1;
        if (f) state[i] = 2;
        else und = 1;
    }
    __shared__ int cnt[4];
    const int wave = static_cast<int>(threadIdx.x) >> 6;
    const unsigned long long b = __ballot(und);
    if ((threadIdx.x & 63) == 0) cnt[wave] = __popcll(b);
    __syncthreads();
    if (threadIdx.x == 0) {
        const int t = cnt[0] + cnt[1] + cnt[2] + cnt[3];
        if (t > 0) atomicAdd(left, t);
    }
}

__global__ void k_flag_state(int n, const unsigned char* __restrict__ state, int* __restrict__ flag) {
    const int i = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (i < n) flag[i] = state[i] == 1;
}

// agg1: roots -> their id, their strong neighbours -> the root's id, other
// rows -1 (two roots are never neighbours of one row).
__global__ void k_agg_join_roots(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                 const unsigned char* __restrict__ strong, const unsigned char* __restrict__ state,
                                 const int* __restrict__ rid, int* __restrict__ agg1) {
    const int i = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (i >= n) return;
    int a = -1;
    if (state[i] == 1) {
        a = rid[i];
    } else {
        for (int e = rp[i]; e < rp[i + 1]; ++e)
            if (strong[e] && state[ci[e]] == 1) a = rid[ci[e]];
    }
    agg1[i] = a;
}

// agg: agg1, or the aggregate of the strongest neighbour with agg1 >= 0
// (largest |a_ij|, then the smallest j); -1 when there is none.
__global__ void k_agg_join_rest(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                const double* __restrict__ va, const unsigned char* __restrict__ strong,
                                const int* __restrict__ agg1, int* __restrict__ agg, int* __restrict__ flag) {
    const int i = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (i >= n) return;
    int a = agg1[i];
    if (a < 0) {
        double best = -1.0;
        int bj = -1;
        for (int e = rp[i]; e < rp[i + 1]; ++e) {
            const int j = ci[e];
            if (!strong[e] || agg1[j] < 0) continue;
            const double v = fabs(va[e]);
            if (v > best || (v == best && j < bj)) {
                best = v;
                bj = j;
            }
        }
        if (bj >= 0) a = agg1[bj];
    }
    agg[i] = a;
    flag[i] = a < 0;
}

// Rows left over: aggregates nroots + (their rank among the left-over rows).
__global__ void k_agg_left(int n, const int* __restrict__ nroots, const int* __restrict__ pos,
                           int* __restrict__ agg) {
    const int i = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (i < n && agg[i] < 0) agg[i] = *nroots + pos[i];
}

// Member lists: cnt[a] = |aggregate a|; after the scan, fill (any order)
// then sort each list ascending.
__global__ void k_agg_count(int n, const int* __restrict__ agg, int* __restrict__ cnt) {
    const int i = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (i < n) atomicAdd(&cnt[agg[i]], 1);
}

__global__ void k_agg_fill(int n, const int* __restrict__ agg, const int* __restrict__ mptr, int* __restrict__ cur,
                           int* __restrict__ mem) {
    const int i = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (i >= n) return;
    const int a = agg[i];
    mem[mptr[a] + atomicAdd(&cur[a], 1)] = i;
}

__global__ void k_agg_sort(int nagg, const int* __restrict__ mptr, int* __restrict__ mem) {
    const int a = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (a >= nagg) return;
    const int b = mptr[a], e = mptr[a + 1];
    for (int k = b + 1; k < e; ++k) {
        const int v = mem[k];
        int j = k;
        while (j > b && mem[j - 1] > v) {
            mem[j] = mem[j - 1];
            --j;
        }
        mem[j] = v;
    }
}

// The coarse rows' member lists in the finer level's solve order: cnt[p] =
// the size of the aggregate of coarse solve row p (perm: solve -> natural,
// nullptr for the natural order), then mpos[mp[p] + k] = the solve position
// of the aggregate's k-th row.
__global__ void k_mg_member_count(int np, const int* __restrict__ perm, const int* __restrict__ mptr,
                                  int* __restrict__ cnt) {
    const int p = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (p >= np) return;
    const int a = perm != nullptr ? perm[p] : p;
    cnt[p] = a < 0 ? 0 : mptr[a + 1] - mptr[a];
}

__global__ void k_mg_member_pos(int np, const int* __restrict__ perm, const int* __restrict__ mptr,
                                const int* __restrict__ mem, const int* __restrict__ inv_fine,
                                const int* __restrict__ mp, int* __restrict__ mpos) {
    const int p = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (p >= np) return;
    const int a = perm != nullptr ? perm[p] : p;
    if (a < 0) return;
    const int b = mptr[a], o = mp[p];
    for (int k = 0; k < mptr[a + 1] - b; ++k) mpos[o + k] = inv_fine[mem[b + k]];
}

// Smoother blocks cut from aggregates (members ascending in mptr / mem):
// aggregate a gives ceil(size / 8) pieces of consecutive members, the first
// size % pieces of them one row longer (sizes as equal as possible), in
// aggregate order from piece poff[a].
__global__ void k_piece_count(int nagg, const int* __restrict__ mptr, int* __restrict__ cnt) {
    const int a = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (a < nagg) cnt[a] = (mptr[a + 1] - mptr[a] + 7) / 8;
}

__global__ void k_piece_fill(int nagg, const int* __restrict__ mptr, const int* __restrict__ mem,
                             const int* __restrict__ poff, int* __restrict__ slots) {
    const int a = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (a >= nagg) return;
    const int b = mptr[a], size = mptr[a + 1] - b;
    const int k = (size + 7) / 8;
    if (k == 0) return;
    const int q = size / k, rem = size % k;
    int m = b;
    for (int p = 0; p < k; ++p) {
        const int len = q + (p < rem ? 1 : 0);
        int* out = slots + static_cast<size_t>(poff[a] + p) * 8;
        for (int j = 0; j < 8; ++j) out[j] = j < len ? mem[m + j] : -1;
        m += len;
    }
}

// Candidate count of coarse row u: the entries of its aggregate's rows.
__global__ void k_gal_count(int units, const int* __restrict__ mptr, const int* __restrict__ mem,
                            const int* __restrict__ rp, int* __restrict__ cnt) {
    const int u = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (u >= units) return;
    int c = 0;
    for (int k = mptr[u]; k < mptr[u + 1]; ++k) c += rp[mem[k] + 1] - rp[mem[k]];
    cnt[u] = c;
}

// Coarse row u of P^T A P: the entries (agg[col], value) of its aggregate's
// rows (ascending) in CSR order, equal coarse columns summed in the order
// they come, the distinct columns sorted ascending and written from off[u];
// their count in dcnt[u].  The distinct list is kept in LDS (kGalCap per
// thread, column-major so the 64 threads of a workgroup touch 64 banks); a
// row with more distinct columns sorts all its candidates in place in
// global memory instead (insertion sort, stable: the same sums in the same
// order).
constexpr int kGalThreads = 64;
constexpr int kGalCap = 32;
__global__ __launch_bounds__(kGalThreads) void k_gal_fill(int units, const int* __restrict__ mptr,
                                                          const int* __restrict__ mem, const int* __restrict__ rp,
                                                          const int* __restrict__ ci, const double* __restrict__ va,
                                                          const int* __restrict__ agg, const int* __restrict__ off,
                                                          int* __restrict__ ccol, double* __restrict__ cval,
                                                          int* __restrict__ dcnt) {
    __shared__ int lcol[kGalCap * kGalThreads];
    __shared__ double lval[kGalCap * kGalThreads];
    const int tid = static_cast<int>(threadIdx.x);
    const int u = static_cast<int>(blockIdx.x) * kGalThreads + tid;
    if (u >= units) return;
    int* col = lcol + tid;
    double* val = lval + tid;
    int d = 0;
    bool over = false;
    for (int k = mptr[u]; k < mptr[u + 1] && !over; ++k) {
        const int row = mem[k];
        for (int e = rp[row]; e < rp[row + 1]; ++e) {
            const int c = agg[ci[e]];
            int j = 0;
            while (j < d && col[j * kGalThreads] != c) ++j;
            if (j < d) {
                val[j * kGalThreads] += va[e];
            } else if (d < kGalCap) {
                col[d * kGalThreads] = c;
                val[d * kGalThreads] = va[e];
                ++d;
            } else {
                over = true;
                break;
            }
        }
    }
    const int o = off[u];
    if (!over) {
        for (int i = 1; i < d; ++i) {
            const int c = col[i * kGalThreads];
            const double v = val[i * kGalThreads];
            int j = i;
            while (j > 0 && col[(j - 1) * kGalThreads] > c) {
                col[j * kGalThreads] = col[(j - 1) * kGalThreads];
                val[j * kGalThreads] = val[(j - 1) * kGalThreads];
                --j;
            }
            col[j * kGalThreads] = c;
            val[j * kGalThreads] = v;
        }
        for (int j = 0; j < d; ++j) {
            ccol[o + j] = col[j * kGalThreads];
            cval[o + j] = val[j * kGalThreads];
        }
        dcnt[u] = d;
        return;
    }
    int m = 0;
    for (int k = mptr[u]; k < mptr[u + 1]; ++k) {
        const int row = mem[k];
        for (int e = rp[row]; e < rp[row + 1]; ++e) {
            const int c = agg[ci[e]];
            const double v = va[e];
            int j = m;
            while (j > 0 && ccol[o + j - 1] > c) {
                ccol[o + j] = ccol[o + j - 1];
                cval[o + j] = cval[o + j - 1];
                --j;
            }
            ccol[o + j] = c;
            cval[o + j] = v;
            ++m;
        }
    }
    int dd = 0;
    for (int j = 0; j < m; ++j) {
        if (dd > 0 && ccol[o + dd - 1] == ccol[o + j]) {
            cval[o + dd - 1] += cval[o + j];
        } else {
            ccol[o + dd] = ccol[o + j];
            cval[o + dd] = cval[o + j];
            ++dd;
        }
    }
    dcnt[u] = dd;
}

__global__ void k_gal_compact(int units, const int* __restrict__ off, const int* __restrict__ crp,
                              const int* __restrict__ ccol, const double* __restrict__ cval, int* __restrict__ cci,
                              double* __restrict__ cva) {
    const int u = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (u >= units) return;
    const int o = off[u], b = crp[u], n = crp[u + 1] - crp[u];
    for (int j = 0; j < n; ++j) {
        cci[b + j] = ccol[o + j];
        cva[b + j] = cval[o + j];
    }
}

// inv[perm[i]] = i for the real solve rows.
__global__ void k_inv_perm(int n_pad, const int* __restrict__ perm, int* __restrict__ inv) {
    const int i = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (i < n_pad && perm[i] >= 0) inv[perm[i]] = i;
}

// The exact inverse of every diagonal 8x8 block of a SELL matrix (FP32
// values, solve order), as full FP32 blocks for k_block_factor: one thread
// per block, Cholesky in FP64; a block that is not positive definite keeps
// its diagonal inverse.
__global__ void k_block_exact_full(SellView A, int nq, float* __restrict__ full) {
    const int q = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (q >= nq) return;
    double a[8][8] = {};
    for (int i = 0; i < 8; ++i) {
        const int row = q * 8 + i;
        const int sl = row / kC, lane = row % kC;
        const int j0 = A.slice_ptr[sl], j1 = A.slice_ptr[sl + 1];
        const int base = A.slice_base[sl];
        for (int j = j0; j < j1; ++j) {
            const size_t e = static_cast<size_t>(j) * kC + static_cast<size_t>(lane);
            const int c = j == j0 ? row : (A.idx16 ? base + static_cast<int>(A.col16[e]) : A.col32[e]);
            if (c >= q * 8 && c < q * 8 + 8) a[i][c - q * 8] += static_cast<double>(A.val[e]);
        }
    }
    double l[8][8] = {};
    bool ok = true;
    for (int k = 0; k < 8 && ok; ++k) {
        double d = a[k][k];
        for (int t = 0; t < k; ++t) d -= l[k][t] * l[k][t];
        if (!(d > 0.0)) {
            ok = false;
            break;
        }
        l[k][k] = sqrt(d);
        for (int i = k + 1; i < 8; ++i) {
            double v = a[i][k];
            for (int t = 0; t < k; ++t) v -= l[i][t] * l[k][t];
            l[i][k] = v / l[k][k];
        }
    }
    float* out = full + static_cast<size_t>(q) * 64;
    if (!ok) {
        for (int i = 0; i < 8; ++i)
            for (int k = 0; k < 8; ++k) out[i * 8 + k] = (i == k && a[i][i] > 0.0) ? static_cast<float>(1.0 / a[i][i]) : (i == k ? 1.0f : 0.0f);
        return;
    }
    // inverse of L (lower), then A^-1 = L^-T L^-1
    double li[8][8] = {};
    for (int k = 0; k < 8; ++k) {
        li[k][k] = 1.0 / l[k][k];
        for (int i = k + 1; i < 8; ++i) {
            double v = 0.0;
            for (int t = k; t < i; ++t) v -= l[i][t] * li[t][k];
            li[i][k] = v / l[i][i];
        }
    }
    for (int i = 0; i < 8; ++i)
        for (int k = 0; k < 8; ++k) {
            double v = 0.0;
            for (int t = (i > k ? i : k); t < 8; ++t) v += li[t][i] * li[t][k];
            out[i * 8 + k] = static_cast<float>(v);
        }
}

// Dense n x n matrix (row-major FP64) of a CSR.
__global__ void k_dense_from_csr(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                 const double* __restrict__ va, double* __restrict__ dense) {
    const int i = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (i >= n) return;
    for (int e = rp[i]; e < rp[i + 1]; ++e) dense[static_cast<size_t>(i) * n + ci[e]] += va[e];
}

// Gauss-Jordan inverse of an SPD matrix (no pivoting): pivot k in one
// kernel that reads the matrix of step k - 1 (a) and writes step k's (b), so
// no entry is read after it changes; a non-positive pivot sets *bad.
// One Gauss-Jordan sweep over the pivot block K = [k, k + bk) (bk <= 8) of
// the n x n matrix a into b: with P = a_KK, b_KK = P^-1, b_Kj = P^-1 a_Kj,
// b_iK = -a_iK P^-1 and b_ij = a_ij - a_iK P^-1 a_Kj.  Sweeping every block
// in turn leaves the inverse (the block form of the one-pivot sweep, which
// it equals in exact arithmetic).  Every workgroup inverts P itself by the
// same sweep, one pivot at a time, one element per lane of wavefront 0
// (FP64); *bad = 1 on a non-positive pivot (P of an SPD matrix's sweep is
// SPD).
__global__ __launch_bounds__(256) void k_gj_block(int n, int k, int bk, const double* __restrict__ a,
                                                  double* __restrict__ b, int* __restrict__ bad) {
    __shared__ double P[8][8];
    const int tid = static_cast<int>(threadIdx.x);
    if (tid < 64) {
        const int r = tid >> 3, c = tid & 7;
        P[r][c] = (r < bk && c < bk) ? a[static_cast<size_t>(k + r) * n + k + c] : (r == c ? 1.0 : 0.0);
    }
    __syncthreads();
    for (int q = 0; q < bk; ++q) {
        const double pq = P[q][q];
        double v = 0.0;
        if (tid < 64) {
            const int r = tid >> 3, c = tid & 7;
            if (r == q && c == q) v = 1.0 / pq;
            else if (r == q) v = P[q][c] / pq;
            else if (c == q) v = -P[r][q] / pq;
            else v = P[r][c] - P[r][q] * P[q][c] / pq;
        }
        __syncthreads();
        if (tid < 64) P[tid >> 3][tid & 7] = v;
        if (tid == 0 && blockIdx.x == 0 && !(pq > 0.0)) *bad = 1;
        __syncthreads();
    }
    const size_t idx = static_cast<size_t>(blockIdx.x) * 256 + static_cast<size_t>(tid);
    if (idx >= static_cast<size_t>(n) * n) return;
    const int i = static_cast<int>(idx / n), j = static_cast<int>(idx % n);
    const bool iK = i >= k && i < k + bk, jK = j >= k && j < k + bk;
    const double* ai = a + static_cast<size_t>(i) * n + k;  // a_iK
    double v;
    if (iK && jK) {
        v = P[i - k][j - k];
    } else if (iK) {
        v = 0.0;
        for (int q = 0; q < bk; ++q) v += P[i - k][q] * a[static_cast<size_t>(k + q) * n + j];
    } else if (jK) {
        v = 0.0;
        for (int q = 0; q < bk; ++q) v -= ai[q] * P[q][j - k];
    } else {
        double akj[8];
        for (int r = 0; r < 8; ++r) akj[r] = r < bk ? a[static_cast<size_t>(k + r) * n + j] : 0.0;
        double s = 0.0;
        for (int q = 0; q < bk; ++q) {
            double t = 0.0;
            for (int r = 0; r < bk; ++r) t += P[q][r] * akj[r];
            s += ai[q] * t;
        }
        v = a[idx] - s;
    }
    b[idx] = v;
}

__global__ void k_f64_to_f32(size_t n, const double* __restrict__ a, float* __restrict__ out) {
    const size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) out[i] = static_cast<float>(a[i]);
}
}  // namespace

// ---------------------------------------------------------------------------
// Learned block preconditioner: features of the aligned 8-row solve-order
// blocks, read straight from the SELL matrix (block_rows = 8 keeps every
// natural 8-row group together as 8 consecutive solve rows, so a solve-order
// block is a natural block).  Same definitions as the host features
// (BlockSet::from_csr): S = D^-1/2 A D^-1/2, the 36 upper entries of S_bb,
// the 8 off-block sums of |s|, 4 zeros.

__device__ __forceinline__ double sell_value(const SellView& A, size_t e) {
    if (A.val64 != nullptr) return A.val64[e];
    return static_cast<double>(A.val[e]) + (A.val_lo != nullptr ? static_cast<double>(A.val_lo[e]) : 0.0);
}

// dis[row] = 1/sqrt(a_rr) from the stored diagonal (slot 0), 1 where it is
// not positive (rows without a diagonal store 0 there).
__global__ void k_inv_sqrt_diag(SellView A, double* __restrict__ dis) {
    const int row = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (row >= A.n_pad) return;
    const int sl = row / kC;
    const double d = sell_value(A, static_cast<size_t>(A.slice_ptr[sl]) * kC + static_cast<size_t>(row % kC));
    dis[row] = d > 0.0 ? 1.0 / sqrt(d) : 1.0;
}

// One thread per solve row: its row of S_bb (register-resident, selected by
// column without dynamic indexing) and its off-block sum.  Thread r of a
// block writes the upper-triangle entries (r, c >= r) of the feature row.
template <bool IDX16>
__global__ void k_block_features(SellView A, const double* __restrict__ dis, float* __restrict__ feat,
                                 float* __restrict__ scale) {
    const int row = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (row >= A.n_pad) return;
    const int sl = row / kC, lane = row % kC;
    const int j0 = A.slice_ptr[sl], w = A.slice_ptr[sl + 1] - j0, base = A.slice_base[sl];
    const double di = dis[row];
    const int blk0 = row & ~7;
    double a[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double off = 0.0;
    for (int k = 0; k < w; ++k) {
        const size_t e = static_cast<size_t>(j0 + k) * kC + static_cast<size_t>(lane);
        const int c = k == 0 ? row : (IDX16 ? base + static_cast<int>(A.col16[e]) : A.col32[e]);
        const double v = sell_value(A, e) * di * dis[c];
        if ((c & ~7) == blk0) {
#pragma unroll
            for (int q = 0; q < 8; ++q) a[q] += (c & 7) == q ? v : 0.0;
        } else {
            off += fabs(v);
        }
    }
    const int r = row & 7;
    float* f = feat + static_cast<size_t>(row >> 3) * 48;
    const int o = 8 * r - r * (r - 1) / 2 - r;  // feature index of (r, c) is o + c
#pragma unroll
    for (int c = 0; c < 8; ++c)
        if (c >= r) f[o + c] = static_cast<float>(a[c]);
    f[36 + r] = static_cast<float>(off);
    if (r < 4) f[44 + r] = 0.0f;
    scale[row] = static_cast<float>(di);
}

// Rows and columns of padding rows (perm < 0) become those of the identity,
// so padding rows stay decoupled (r = z = 0 there).  One thread per solve row
// rewrites its own block row.
__global__ void k_mask_padding_blocks(int n_pad, const int* __restrict__ perm, float* __restrict__ blocks) {
    const int row = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (row >= n_pad) return;
    const int blk0 = row & ~7, j = row & 7;
    const bool pad_row = perm[row] < 0;
    float* b = blocks + static_cast<size_t>(row) * 8;
#pragma unroll
    for (int c = 0; c < 8; ++c)
        if (pad_row || perm[blk0 + c] < 0) b[c] = c == j ? 1.0f : 0.0f;
}

// One workgroup-grid launch on L.stream, timed by L's events when set.
#define XP_LAUNCH(L, K, G, ...)                                                                   \
    do {                                                                                       \
        if ((L).ev_start != nullptr)                                                           \
            hipExtLaunchKernelGGL(K, G, dim3(kBlock), 0, (L).stream, (L).ev_start, (L).ev_stop, 0, \
                                  __VA_ARGS__);                                                \
        else                                                                                   \
            hipLaunchKernelGGL(K, G, dim3(kBlock), 0, (L).stream, __VA_ARGS__);                 \
    } while (0)

// SELL kernels templated on <IDX16, V>.
#define XP_LAUNCH_SELL(L, A, K, ...)                                                              \
    do {                                                                                       \
        if ((A).idx16) {                                                                       \
            if ((A).variant == 2) XP_LAUNCH(L, (K<true, 2>), dim3((L).grid), __VA_ARGS__);       \
            else if ((A).variant == 1) XP_LAUNCH(L, (K<true, 1>), dim3((L).grid), __VA_ARGS__);  \
            else XP_LAUNCH(L, (K<true, 0>), dim3((L).grid), __VA_ARGS__);                        \
        } else {                                                                               \
            if ((A).variant == 2) XP_LAUNCH(L, (K<false, 2>), dim3((L).grid), __VA_ARGS__);      \
            else if ((A).variant == 1) XP_LAUNCH(L, (K<false, 1>), dim3((L).grid), __VA_ARGS__); \
            else XP_LAUNCH(L, (K<false, 0>), dim3((L).grid), __VA_ARGS__);                       \
        }                                                                                      \
    } while (0)

#define XP_DISPATCH_SELL(A, KERNEL, ...)                                                            \
    do {                                                                                        \
        if ((A).idx16) {                                                                        \
            if ((A).variant == 1)                                                               \
                hipLaunchKernelGGL((KERNEL<true, 1>), __VA_ARGS__);                             \
            else                                                                                \
                hipLaunchKernelGGL((KERNEL<true, 0>), __VA_ARGS__);                             \
        } else {                                                                                \
            if ((A).variant == 1)                                                               \
                hipLaunchKernelGGL((KERNEL<false, 1>), __VA_ARGS__);                            \
            else                                                                                \
                hipLaunchKernelGGL((KERNEL<false, 0>), __VA_ARGS__);                            \
        }                                                                                       \
    } while (0)

void launch_pcg_init(const Launch& L, const SellView& A, SysState* st, const float* x,
                     const float* b, const float* minv, const BlockPre* blockM, float* r, float* z, float* p) {
    XP_DISPATCH_SELL(A, k_pcg_init, dim3(L.grid), dim3(kBlock), 0, L.stream, L, A, x, b, minv, blockM, r, z, p);
    hipLaunchKernelGGL(k_pcg_init_state, dim3(A.nsys), dim3(kBlock), 0, L.stream, L, st);
}

void launch_pcg_spmv(const Launch& L, const SellView& A, const SysState* st, const float* p, float* ap) {
    XP_LAUNCH_SELL(L, A, k_pcg_spmv, L, A, st, p, ap);
}

void launch_pcg_spmv_fused(const Launch& L, const SellView& A, const SysState* st_cur, SysState* st_next,
                           TraceRow* trace, int trace_stride, const float* z, float* p, float* ap) {
    XP_LAUNCH_SELL(L, A, k_pcg_spmv_fused, L, A, st_cur, st_next, trace, trace_stride, z, p, ap);
}

void launch_pcg_update_xrz(const Launch& L, const SysState* st, float* x, const float* p, float* r,
                           const float* ap, const float* minv, const BlockPre* blockM, float* z) {
    if (blockM != nullptr)
        XP_LAUNCH(L, k_pcg_update_xrz<true>, dim3(L.grid), L, st, x, p, r, ap, minv, blockM, z);
    else
        XP_LAUNCH(L, k_pcg_update_xrz<false>, dim3(L.grid), L, st, x, p, r, ap, minv, blockM, z);
}

void launch_pcg_update_p(const Launch& L, const SysState* st_cur, SysState* st_next, TraceRow* trace,
                         int trace_stride, const float* z, float* p) {
    XP_LAUNCH(L, k_pcg_update_p, dim3(L.grid), L, st_cur, st_next, trace, trace_stride, z, p);
}

void launch_stream(int grid, const void* src, void* dst, size_t bytes, unsigned* sink, hipStream_t s,
                   hipEvent_t e0, hipEvent_t e1, int nontemporal) {
    const size_t n16 = bytes / 16;
    if (dst != nullptr && nontemporal != 0) {
        if (nontemporal == 2)
            hipExtLaunchKernelGGL(k_stream_copy_nt<true>, dim3(grid), dim3(kBlock), 0, s, e0, e1, 0,
                                  static_cast<const uint4*>(src), static_cast<uint4*>(dst), n16);
        else
            hipExtLaunchKernelGGL(k_stream_copy_nt<false>, dim3(grid), dim3(kBlock), 0, s, e0, e1, 0,
                                  static_cast<const uint4*>(src), static_cast<uint4*>(dst), n16);
    } else if (dst == nullptr && nontemporal != 0)
        hipExtLaunchKernelGGL(k_stream_read_nt, dim3(grid), dim3(kBlock), 0, s, e0, e1, 0,
                              static_cast<const uint4*>(src), n16, sink);
    else if (dst == nullptr)
        hipExtLaunchKernelGGL(k_stream_read, dim3(grid), dim3(kBlock), 0, s, e0, e1, 0,
                              static_cast<const uint4*>(src), n16, sink);
    else
        hipExtLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(kBlock), 0, s, e0, e1, 0,
                              static_cast<const uint4*>(src), static_cast<uint4*>(dst), n16);
}

void launch_spmv_f32(const Launch& L, const SellView& A, const float* x, float* y) {
    XP_LAUNCH_SELL(L, A, k_spmv_f32, L, A, x, y);
}

void launch_spmv_f64(const Launch& L, const SellView& A, bool use_lo, const double* x, double* y) {
    if (A.idx16) {
        if (use_lo)
            hipLaunchKernelGGL((k_spmv_f64<true, 1>), dim3(L.grid), dim3(kBlock), 0, L.stream, L, A, x, y);
        else
            hipLaunchKernelGGL((k_spmv_f64<true, 0>), dim3(L.grid), dim3(kBlock), 0, L.stream, L, A, x, y);
    } else {
        if (use_lo)
            hipLaunchKernelGGL((k_spmv_f64<false, 1>), dim3(L.grid), dim3(kBlock), 0, L.stream, L, A, x, y);
        else
            hipLaunchKernelGGL((k_spmv_f64<false, 0>), dim3(L.grid), dim3(kBlock), 0, L.stream, L, A, x, y);
    }
}

void launch_residual_f64(const Launch& L, const SellView& A, bool use_lo, const double* x,
                         const double* b, double* r, float* r32, double* out_rr) {
    if (A.idx16) {
        if (use_lo)
            hipLaunchKernelGGL((k_residual_f64<true, 1>), dim3(L.grid), dim3(kBlock), 0, L.stream, L, A, x, b, r, r32);
        else
            hipLaunchKernelGGL((k_residual_f64<true, 0>), dim3(L.grid), dim3(kBlock), 0, L.stream, L, A, x, b, r, r32);
    } else {
        if (use_lo)
            hipLaunchKernelGGL((k_residual_f64<false, 1>), dim3(L.grid), dim3(kBlock), 0, L.stream, L, A, x, b, r, r32);
        else
            hipLaunchKernelGGL((k_residual_f64<false, 0>), dim3(L.grid), dim3(kBlock), 0, L.stream, L, A, x, b, r, r32);
    }
    hipLaunchKernelGGL(k_finish_aux, dim3(A.nsys), dim3(kBlock), 0, L.stream, L, out_rr);
}

void launch_axpy_f64_f32(int n_pad, const float* d, double* x, hipStream_t s) {
    if (n_pad <= 0) return;
    hipLaunchKernelGGL(k_axpy_f64_f32, dim3(blocks_for(static_cast<size_t>(n_pad), 256)), dim3(256), 0, s, n_pad, d, x);
}

void launch_init_pcg_f64(int n, const double* b, const double* ax, const double* m_inv,
                         double* r, double* z, double* p, double* out, double* partials, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_init_pcg_f64, dim3(grid), dim3(kBlock), 0, s, n, b, ax, m_inv, r, z, p, partials);
    hipLaunchKernelGGL(k_flat_finish<2>, dim3(1), dim3(kBlock), 0, s, partials, grid, out);
}

void launch_dot_f64(int n, const double* a, const double* b, double* out, double* partials, int grid,
                    hipStream_t s) {
    hipLaunchKernelGGL(k_dot_f64, dim3(grid), dim3(kBlock), 0, s, n, a, b, partials);
    hipLaunchKernelGGL(k_flat_finish<1>, dim3(1), dim3(kBlock), 0, s, partials, grid, out);
}

void launch_update_xrz_f64(int n, double* x, const double* p, double* r, const double* ap,
                           const double* m_inv, double* z, double alpha, double* out,
                           double* partials, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_update_xrz_f64, dim3(grid), dim3(kBlock), 0, s, n, x, p, r, ap, m_inv, z, alpha, partials);
    hipLaunchKernelGGL(k_flat_finish<2>, dim3(1), dim3(kBlock), 0, s, partials, grid, out);
}

void launch_update_p_f64(int n, const double* z, double* p, double beta, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_update_p_f64, dim3(blocks_for(static_cast<size_t>(n), 256)), dim3(256), 0, s, n, z, p, beta);
}

#define XP_PERM_LAUNCH(name, kern, S, D)                                                          \
    void name(int n_pad, const int* perm, const S* src, D* dst, hipStream_t s) {                  \
        if (n_pad <= 0) return;                                                                   \
        hipLaunchKernelGGL((kern<S, D>), dim3(blocks_for(static_cast<size_t>(n_pad), 256)), dim3(256), 0, s, \
                           n_pad, perm, src, dst);                                                \
    }
XP_PERM_LAUNCH(launch_gather_d2f, k_gather, double, float)
XP_PERM_LAUNCH(launch_gather_d2d, k_gather, double, double)
XP_PERM_LAUNCH(launch_gather_f2f, k_gather, float, float)
XP_PERM_LAUNCH(launch_scatter_f2d, k_scatter, float, double)
XP_PERM_LAUNCH(launch_scatter_d2d, k_scatter, double, double)
XP_PERM_LAUNCH(launch_scatter_f2f, k_scatter, float, float)
#undef XP_PERM_LAUNCH

void launch_convert_f2d(int n, const float* src, double* dst, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_convert_f2d, dim3(blocks_for(static_cast<size_t>(n), 256)), dim3(256), 0, s, n, src, dst);
}

void launch_fill_f32(size_t n, float v, float* dst, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_fill<float>, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, v, dst);
}

void launch_fill_u32(size_t n, unsigned v, unsigned* dst, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_fill<unsigned>, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, v, dst);
}

void launch_block_features(const SellView& A, double* dis, float* feat, float* scale, hipStream_t s) {
    const int g = static_cast<int>(blocks_for(static_cast<size_t>(A.n_pad), 256));
    hipLaunchKernelGGL(k_inv_sqrt_diag, dim3(g), dim3(256), 0, s, A, dis);
    if (A.idx16)
        hipLaunchKernelGGL(k_block_features<true>, dim3(g), dim3(256), 0, s, A, dis, feat, scale);
    else
        hipLaunchKernelGGL(k_block_features<false>, dim3(g), dim3(256), 0, s, A, dis, feat, scale);
}

void launch_pcg_replace(const Launch& L, const SellView& A, bool use_lo, const double* x, const double* b,
                        const float* minv, const BlockPre* blockM, float* r, float* z, double* out,
                        const MgPlan* mg) {
    const dim3 g(L.grid), blk(kBlock);
    if (A.idx16) {
        if (use_lo)
            hipLaunchKernelGGL((k_pcg_replace<true, 1>), g, blk, 0, L.stream, L, A, x, b, minv, blockM, r, z);
        else
            hipLaunchKernelGGL((k_pcg_replace<true, 0>), g, blk, 0, L.stream, L, A, x, b, minv, blockM, r, z);
    } else {
        if (use_lo)
            hipLaunchKernelGGL((k_pcg_replace<false, 1>), g, blk, 0, L.stream, L, A, x, b, minv, blockM, r, z);
        else
            hipLaunchKernelGGL((k_pcg_replace<false, 0>), g, blk, 0, L.stream, L, A, x, b, minv, blockM, r, z);
    }
    if (mg != nullptr) launch_vcycle(*mg, L, nullptr, r, z);
    hipLaunchKernelGGL(k_replace_sums, dim3(A.nsys), blk, 0, L.stream, L, out);
}

#define XP_DISPATCH_ACC(A, use_lo, KERNEL, ...)                                  \
    do {                                                                     \
        if ((A).idx16) {                                                     \
            if (use_lo)                                                      \
                hipLaunchKernelGGL((KERNEL<true, 1>), __VA_ARGS__);          \
            else                                                             \
                hipLaunchKernelGGL((KERNEL<true, 0>), __VA_ARGS__);          \
        } else {                                                             \
            if (use_lo)                                                      \
                hipLaunchKernelGGL((KERNEL<false, 1>), __VA_ARGS__);         \
            else                                                             \
                hipLaunchKernelGGL((KERNEL<false, 0>), __VA_ARGS__);         \
        }                                                                    \
    } while (0)

void launch_pcg_init_acc(const Launch& L, const SellView& A, bool use_lo, SysState* st, const double* x,
                         const double* b, const float* minv, const BlockPre* blockM, float* r, float* z,
                         const MgPlan* mg) {
    XP_DISPATCH_ACC(A, use_lo, k_pcg_replace, dim3(L.grid), dim3(kBlock), 0, L.stream, L, A, x, b, minv, blockM, r, z);
    if (mg != nullptr) launch_vcycle(*mg, L, nullptr, r, z);
    hipLaunchKernelGGL(k_pcg_init_state, dim3(A.nsys), dim3(kBlock), 0, L.stream, L, st);
}

void launch_pcg_spmv_fused_acc(const Launch& L, const SellView& A, int vmode, const SysState* st_cur,
                               SysState* st_next, TraceRow* trace, int trace_stride, const float* z,
                               const double* p_old, double* p_new, double* ap) {
#define XP_FUSED_ACC(I, V)                                                                                   \
    do {                                                                                                     \
        if (A.variant == 2)                                                                                  \
            XP_LAUNCH(L, (k_pcg_spmv_fused_acc<I, V, true>), dim3(L.grid), L, A, st_cur, st_next, trace,     \
                      trace_stride, z, p_old, p_new, ap);                                                    \
        else                                                                                                 \
            XP_LAUNCH(L, (k_pcg_spmv_fused_acc<I, V, false>), dim3(L.grid), L, A, st_cur, st_next, trace,    \
                      trace_stride, z, p_old, p_new, ap);                                                    \
    } while (0)
    if (A.idx16) {
        if (vmode == 2) XP_FUSED_ACC(true, 2);
        else if (vmode == 1) XP_FUSED_ACC(true, 1);
        else XP_FUSED_ACC(true, 0);
    } else {
        if (vmode == 2) XP_FUSED_ACC(false, 2);
        else if (vmode == 1) XP_FUSED_ACC(false, 1);
        else XP_FUSED_ACC(false, 0);
    }
#undef XP_FUSED_ACC
}

void launch_pcg_update_xrz_acc(const Launch& L, const SysState* st, bool odd, double* x, const double* p,
                               const double* p_prev, float* r, const double* ap, const float* minv,
                               const BlockPre* blockM, float* z) {
#define XP_UPD_ACC(B, XM) \
    XP_LAUNCH(L, (k_pcg_update_xrz_acc<B, XM>), dim3(L.grid), L, st, x, p, p_prev, r, ap, minv, blockM, z)
    if (blockM != nullptr) {
        if (odd) XP_UPD_ACC(true, 1);
        else XP_UPD_ACC(true, 0);
    } else {
        if (odd) XP_UPD_ACC(false, 1);
        else XP_UPD_ACC(false, 0);
    }
#undef XP_UPD_ACC
}

void launch_pcg_flush_x_acc(const Launch& L, int nsys, SysState* st, const double* p, double* x) {
    hipLaunchKernelGGL(k_pcg_flush_x_acc, dim3(L.grid), dim3(kBlock), 0, L.stream, L, st, p, x);
    hipLaunchKernelGGL(k_pcg_clear_x_pend, dim3((nsys + 255) / 256), dim3(256), 0, L.stream, nsys, st);
}

void launch_fold_f32_into_f64(int n_pad, float* d, double* x, hipStream_t s) {
    hipLaunchKernelGGL(k_fold_f32_into_f64, dim3(blocks_for(static_cast<size_t>(n_pad), 256)), dim3(256), 0, s, n_pad,
                       d, x);
}

void launch_block_factor(int nblocks, const float* full, BlockPre* out, hipStream_t s) {
    if (nblocks <= 0) return;
    hipLaunchKernelGGL(k_block_factor, dim3(blocks_for(static_cast<size_t>(nblocks), 256)), dim3(256), 0, s, nblocks,
                       full, out);
}

void launch_mask_padding_blocks(int n_pad, const int* perm, float* blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_mask_padding_blocks, dim3(blocks_for(static_cast<size_t>(n_pad), 256)), dim3(256), 0, s, n_pad,
                       perm, blocks);
}


void launch_jacobi_iter(const Launch& L, const SellView& A, const JacState* st_cur, JacState* st_next,
                        const float* b, const float* dinv, const float* x_old, float* x_new, int slot_read,
                        int slot_write) {
    // Slices up to 5 columns (2D 5-point stencils) go two at a time (16-bit
    // columns; one with 32-bit columns); up to
    // 8 or 10 columns one at a time with all loads ahead (more would not
    // fit the 64 registers of full occupancy); wider ones through the loop.
#define XP_JACOBI_LAUNCH(I16, FW, U) \
    XP_LAUNCH(L, (k_jacobi_iter<I16, FW, U>), dim3(L.grid), L, A, st_cur, st_next, b, dinv, x_old, x_new, \
              slot_read, slot_write)
#define XP_JACOBI_WIDTHS(I16)                                  \
    do {                                                       \
        if (A.max_width <= 5) XP_JACOBI_LAUNCH(I16, 5, (I16 ? 2 : 1)); \
        else if (A.max_width <= 8) XP_JACOBI_LAUNCH(I16, 8, 1); \
        else if (A.max_width <= 10) XP_JACOBI_LAUNCH(I16, 10, 1); \
        else XP_JACOBI_LAUNCH(I16, 0, 1);                      \
    } while (0)
    if (A.idx16) XP_JACOBI_WIDTHS(true);
    else XP_JACOBI_WIDTHS(false);
#undef XP_JACOBI_WIDTHS
#undef XP_JACOBI_LAUNCH
}

void launch_jacobi_sweep(const Launch& L, const SellView& A, const float* b, const float* dinv,
                         const float* x_old, float* x_new) {
#define XP_JACOBI_SWEEP(I16, FW, U) \
    XP_LAUNCH(L, (k_jacobi_sweep<I16, FW, U>), dim3(L.grid), L, A, b, dinv, x_old, x_new)
#define XP_JACOBI_SWEEP_WIDTHS(I16)                                      \
    do {                                                                 \
        if (A.max_width <= 5) XP_JACOBI_SWEEP(I16, 5, (I16 ? 2 : 1));    \
        else if (A.max_width <= 8) XP_JACOBI_SWEEP(I16, 8, 1);           \
        else if (A.max_width <= 10) XP_JACOBI_SWEEP(I16, 10, 1);         \
        else XP_JACOBI_SWEEP(I16, 0, 1);                                 \
    } while (0)
    if (A.idx16) XP_JACOBI_SWEEP_WIDTHS(true);
    else XP_JACOBI_SWEEP_WIDTHS(false);
#undef XP_JACOBI_SWEEP_WIDTHS
#undef XP_JACOBI_SWEEP
}

void launch_jacobi_finish(const Launch& L, int nsys, JacState* st, int iters) {
    hipLaunchKernelGGL(k_jacobi_finish, dim3(nsys), dim3(kBlock), 0, L.stream, L, st, iters);
}

void launch_op_init(const Launch& L, int nsys, const float* b, const float* ax, const float* minv, float* r,
                    float* z, float* p, double* metrics) {
    XP_LAUNCH(L, k_op_init, dim3(L.grid), L, b, ax, minv, r, z, p);
    hipLaunchKernelGGL(k_op_finish, dim3(nsys), dim3(kBlock), 0, L.stream, L, static_cast<int>(kSlotRz), 2, metrics);
}

void launch_op_dot(const Launch& L, int nsys, const float* a, const float* b, double* out) {
    XP_LAUNCH(L, k_op_dot, dim3(L.grid), L, a, b);
    hipLaunchKernelGGL(k_op_finish, dim3(nsys), dim3(kBlock), 0, L.stream, L, static_cast<int>(kSlotAux), 1, out);
}

void launch_op_update_xrz(const Launch& L, int nsys, float* x, const float* p, float* r, const float* ap,
                          const float* minv, float* z, double* metrics, double alpha) {
    XP_LAUNCH(L, k_op_update_xrz, dim3(L.grid), L, x, p, r, ap, minv, z, alpha);
    hipLaunchKernelGGL(k_op_finish, dim3(nsys), dim3(kBlock), 0, L.stream, L, static_cast<int>(kSlotRz), 2, metrics);
}

void launch_op_update_p(const Launch& L, const float* z, float* p, double beta) {
    XP_LAUNCH(L, k_op_update_p, dim3(L.grid), L, z, p, beta);
}


// ---- multilevel preconditioner ----
void launch_vcycle(const MgPlan& P, const Launch& L0, const SysState* st, const float* r0, float* z0) {
    const MgLevelView& f = P.lv[0];
#define XP_MG_SELL(LL, A, K, ...)                                                            \
    do {                                                                                     \
        if ((A).idx16) {                                                                     \
            if ((A).variant >= 1) hipLaunchKernelGGL((K<true, 1>), dim3((LL).grid), dim3(kBlock), 0, (LL).stream, __VA_ARGS__); \
            else hipLaunchKernelGGL((K<true, 0>), dim3((LL).grid), dim3(kBlock), 0, (LL).stream, __VA_ARGS__); \
        } else {                                                                             \
            if ((A).variant >= 1) hipLaunchKernelGGL((K<false, 1>), dim3((LL).grid), dim3(kBlock), 0, (LL).stream, __VA_ARGS__); \
            else hipLaunchKernelGGL((K<false, 0>), dim3((LL).grid), dim3(kBlock), 0, (LL).stream, __VA_ARGS__); \
        }                                                                                    \
    } while (0)
    // level 0: z0 = M r from the update kernel; the pre-smoothed approximation is omega z0
    // the restriction writes no partial sums: it may take its own workgroup
    // count (one system: the uniform split needs no work map)
    Launch Lr = L0;
    if (P.grid0 > 0 && L0.u_blocks > 0) {
        Lr.grid = P.grid0;
        Lr.u_blocks = P.grid0;
    }
    XP_MG_SELL(Lr, f.A, k_mg_restrict, Lr, f.A, st, r0, z0, P.omega, f.w);
    // a workgroup with at most one slice shares it among its 16 wavefronts
    auto split = [](const MgLevelView& v) { return v.L.grid >= v.A.nslices; };
    for (int l = 1; l < P.nlev; ++l) {
        const MgLevelView& v = P.lv[l];
        hipLaunchKernelGGL(k_mg_smooth_gather, dim3(v.L.grid), dim3(kBlock), 0, v.L.stream, v.L, st, v.M, v.mp,
                           v.mpos, P.lv[l - 1].w, v.r, P.omega, v.z);
        if (split(v)) {
            if (v.A.idx16)
                hipLaunchKernelGGL(k_mg_restrict_split<true>, dim3(v.L.grid), dim3(kBlock), 0, v.L.stream, v.L, v.A,
                                   st, v.r, v.z, 1.0f, v.w);
            else
                hipLaunchKernelGGL(k_mg_restrict_split<false>, dim3(v.L.grid), dim3(kBlock), 0, v.L.stream, v.L, v.A,
                                   st, v.r, v.z, 1.0f, v.w);
        } else {
            XP_MG_SELL(v.L, v.A, k_mg_restrict, v.L, v.A, st, v.r, v.z, 1.0f, v.w);
        }
    }
    // the prolongation into level l is done by the kernel that produces
    // level l + 1's correction (the coarse apply, then each post-smoothing)
    auto up_into = [&](int l) {
        MgUp pu;
        pu.t = P.lv[l].t;
        pu.z = l == 0 ? z0 : P.lv[l].z;
        pu.cz = l == 0 ? P.omega : 1.0f;
        return pu;
    };
    {
        MgUp pu = up_into(P.nlev - 1);
        pu.mp = P.cmp;
        pu.mpos = P.cmpos;
        hipLaunchKernelGGL(k_mg_coarse, dim3((P.nc + kWaves - 1) / kWaves), dim3(kBlock), 0, L0.stream, st, P.nc,
                           P.cinv, P.lv[P.nlev - 1].w, P.ec, pu, P.gamma);
    }
    for (int l = P.nlev - 1; l >= 1; --l) {
        const MgLevelView& v = P.lv[l];
        MgUp pu = up_into(l - 1);
        pu.mp = v.mp;
        pu.mpos = v.mpos;
        if (split(v)) {
            if (v.A.idx16)
                hipLaunchKernelGGL(k_mg_post_split<true>, dim3(v.L.grid), dim3(kBlock), 0, v.L.stream, v.L, v.A, st,
                                   v.M, v.r, v.t, P.omega, v.z, pu, P.gamma);
            else
                hipLaunchKernelGGL(k_mg_post_split<false>, dim3(v.L.grid), dim3(kBlock), 0, v.L.stream, v.L, v.A, st,
                                   v.M, v.r, v.t, P.omega, v.z, pu, P.gamma);
        } else {
            XP_MG_SELL(v.L, v.A, k_mg_post_norz, v.L, v.A, st, v.M, v.r, v.t, P.omega, v.z, pu, P.gamma);
        }
    }
    XP_MG_SELL(L0, f.A, k_mg_post_rz, L0, f.A, st, f.M, r0, f.t, P.omega, z0);
#undef XP_MG_SELL
}

void launch_mg_block_apply(const Launch& L, const BlockPre* M, const float* r, float* z) {
    hipLaunchKernelGGL(k_mg_smooth, dim3(L.grid), dim3(kBlock), 0, L.stream, L, nullptr, M, r, 1.0f, z);
}

namespace {
inline dim3 rows256(int n) { return dim3(blocks_for(static_cast<size_t>(n > 0 ? n : 1), 256)); }
}  // namespace

void launch_mg_absdiag(int n, const int* rp, const int* ci, const double* va, double* d, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_row_absdiag, rows256(n), dim3(256), 0, s, n, rp, ci, va, d);
}

void launch_mg_strength(int n, const int* rp, const int* ci, const double* va, const double* d, double theta,
                        unsigned char* strong, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_agg_strength, rows256(n), dim3(256), 0, s, n, rp, ci, va, d, theta * theta, strong);
}

void launch_mg_mis_round(int n, const int* rp, const int* ci, const unsigned char* strong, int salt,
                         unsigned char* state, unsigned long long* m1, unsigned char* near, int* left, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_mis_max1, rows256(n), dim3(256), 0, s, n, rp, ci, strong, state, salt, m1);
    hipLaunchKernelGGL(k_mis_root, rows256(n), dim3(256), 0, s, n, rp, ci, strong, m1, salt, state);
    hipLaunchKernelGGL(k_mis_near, rows256(n), dim3(256), 0, s, n, rp, ci, strong, state, near);
    hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(64), 0, s, static_cast<size_t>(1), 0, left);
    hipLaunchKernelGGL(k_mis_exclude, rows256(n), dim3(256), 0, s, n, rp, ci, strong, near, state, left);
}

void launch_mg_flag_roots(int n, const unsigned char* state, int* flag, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_flag_state, rows256(n), dim3(256), 0, s, n, state, flag);
}

void launch_mg_join(int n, const int* rp, const int* ci, const double* va, const unsigned char* strong,
                    const unsigned char* state, const int* rid, int* agg1, int* agg, int* flag, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_agg_join_roots, rows256(n), dim3(256), 0, s, n, rp, ci, strong, state, rid, agg1);
    hipLaunchKernelGGL(k_agg_join_rest, rows256(n), dim3(256), 0, s, n, rp, ci, va, strong, agg1, agg, flag);
}

void launch_mg_left(int n, const int* nroots, const int* pos, int* agg, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_agg_left, rows256(n), dim3(256), 0, s, n, nroots, pos, agg);
}

void launch_mg_member_count(int n, const int* agg, int nagg, int* cnt, hipStream_t s) {
    hipLaunchKernelGGL(k_fill<int>, rows256(nagg), dim3(256), 0, s, static_cast<size_t>(nagg), 0, cnt);
    if (n > 0) hipLaunchKernelGGL(k_agg_count, rows256(n), dim3(256), 0, s, n, agg, cnt);
}

void launch_mg_member_fill(int n, const int* agg, int nagg, const int* mptr, int* cur, int* mem, hipStream_t s) {
    hipLaunchKernelGGL(k_fill<int>, rows256(nagg), dim3(256), 0, s, static_cast<size_t>(nagg), 0, cur);
    if (n > 0) hipLaunchKernelGGL(k_agg_fill, rows256(n), dim3(256), 0, s, n, agg, mptr, cur, mem);
    if (nagg > 0) hipLaunchKernelGGL(k_agg_sort, rows256(nagg), dim3(256), 0, s, nagg, mptr, mem);
}

void launch_mg_member_solve_count(int np, const int* perm, const int* mptr, int* cnt, hipStream_t s) {
    if (np > 0) hipLaunchKernelGGL(k_mg_member_count, rows256(np), dim3(256), 0, s, np, perm, mptr, cnt);
}

void launch_mg_member_solve_pos(int np, const int* perm, const int* mptr, const int* mem, const int* inv_fine,
                                const int* mp, int* mpos, hipStream_t s) {
    if (np > 0)
        hipLaunchKernelGGL(k_mg_member_pos, rows256(np), dim3(256), 0, s, np, perm, mptr, mem, inv_fine, mp, mpos);
}

void launch_mg_piece_count(int nagg, const int* mptr, int* cnt, hipStream_t s) {
    if (nagg > 0) hipLaunchKernelGGL(k_piece_count, rows256(nagg), dim3(256), 0, s, nagg, mptr, cnt);
}

void launch_mg_piece_fill(int nagg, const int* mptr, const int* mem, const int* poff, int* slots, hipStream_t s) {
    if (nagg > 0) hipLaunchKernelGGL(k_piece_fill, rows256(nagg), dim3(256), 0, s, nagg, mptr, mem, poff, slots);
}

void launch_mg_galerkin_count(int units, const int* mptr, const int* mem, const int* rp, int* cnt, hipStream_t s) {
    if (units > 0)
        hipLaunchKernelGGL(k_gal_count, rows256(units), dim3(256), 0, s, units, mptr, mem, rp, cnt);
}

void launch_mg_galerkin_fill(int units, const int* mptr, const int* mem, const int* rp, const int* ci,
                             const double* va, const int* agg, const int* off, int* ccol, double* cval, int* dcnt,
                             hipStream_t s) {
    if (units > 0)
        hipLaunchKernelGGL(k_gal_fill, dim3(blocks_for(static_cast<size_t>(units), kGalThreads)), dim3(kGalThreads),
                           0, s, units, mptr, mem, rp, ci, va, agg, off, ccol, cval, dcnt);
}

void launch_mg_galerkin_compact(int units, const int* off, const int* crp, const int* ccol, const double* cval,
                                int* cci, double* cva, hipStream_t s) {
    if (units > 0)
        hipLaunchKernelGGL(k_gal_compact, rows256(units), dim3(256), 0, s, units, off, crp, ccol, cval, cci, cva);
}

void launch_mg_inv_perm(int n_pad, const int* perm, int* inv, hipStream_t s) {
    hipLaunchKernelGGL(k_inv_perm, rows256(n_pad), dim3(256), 0, s, n_pad, perm, inv);
}

void launch_block_exact_full(const SellView& A, int nq, float* full, hipStream_t s) {
    hipLaunchKernelGGL(k_block_exact_full, dim3(blocks_for(static_cast<size_t>(nq), 256)), dim3(256), 0, s, A, nq,
                       full);
}

void launch_mg_dense_inverse(int n, const int* rp, const int* ci, const double* va, double* work, float* out,
                             int* bad, hipStream_t s) {
    const size_t nn = static_cast<size_t>(n) * static_cast<size_t>(n);
    double* buf[2] = {work, work + nn};
    hipLaunchKernelGGL(k_fill<double>, dim3(blocks_for(nn, 256)), dim3(256), 0, s, nn, 0.0, work);
    hipLaunchKernelGGL(k_dense_from_csr, dim3(blocks_for(static_cast<size_t>(n), 256)), dim3(256), 0, s, n, rp, ci,
                       va, work);
    int steps = 0;
    for (int k = 0; k < n; k += 8, ++steps)
        hipLaunchKernelGGL(k_gj_block, dim3(blocks_for(nn, 256)), dim3(256), 0, s, n, k, std::min(8, n - k),
                           buf[steps & 1], buf[(steps + 1) & 1], bad);
    hipLaunchKernelGGL(k_f64_to_f32, dim3(blocks_for(nn, 256)), dim3(256), 0, s, nn, buf[steps & 1], out);
}

}  // namespace xp

// gfx950 kernels of the exhaustive epistasis search.
//
// Reference semantics (UDC-GAC/fiuncho):
//   combine             src/base/gt_combine.cpp:3-18
//   combine_and_popcnt  src/base/gt_popcnt.cpp:4-23
//   MI                  include/fiuncho/algorithms/MutualInformation.h:58-65,
//                       src/base/mi_compute.cpp:14-41
//   search loop         include/fiuncho/ThreadedSearch.h:142-182
//   distribution        include/fiuncho/Distribution.h:54-176
//
// Scan kernels map one wavefront to one or a few (k-1)-SNP prefixes and its
// 64 lanes to 64 consecutive candidates for the last SNP (a tile).  The
// prefix rows are wave-uniform 16-byte buffer loads, the lanes' rows one
// coalesced 1 KiB line per row and quad; every VALU instruction of the inner
// loop is a v_and_b32 / v_bitop3_b32 or an accumulating v_bcnt_u32_b32.
// Counts stay in VGPRs; MI and the top-N test run in the same kernel (no
// count table is written).  The scans are written as per-wavefront units so
// that one-scan launches and the persistent batched launch share them.
#include "fiu_kernels.h"

#include <math.h>

#include <algorithm>
#include <type_traits>
#include <utility>

namespace fiu {

template <int N> struct Pow3 { static constexpr int v = 3 * Pow3<N - 1>::v; };
template <> struct Pow3<0> { static constexpr int v = 1; };

__device__ __forceinline__ uint32_t comp(const uint4 &v, int w)
{
    return w == 0 ? v.x : w == 1 ? v.y : w == 2 ? v.z : v.w;
}

// Buffer resource over the three genotype rows of one quad (all SNPs of one
// class); loads through it are in-order vector loads whose offset can live in
// an SGPR (wave-uniform prefix rows) or a VGPR (the lanes' own SNPs), with no
// 64-bit address arithmetic, and reads past the rows return zero.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t quad_rsrc(const uint4 *row, int np)
{
    return __builtin_amdgcn_make_buffer_rsrc((void *)row, (short)0, 3 * np * 16, 0x00020000);
}

// Row z of SNP s (wave-uniform s): 16 bytes broadcast to every lane.
__device__ __forceinline__ uint4 ld_uniform(__amdgpu_buffer_rsrc_t r, int np, int z, int s)
{
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, 0, (z * np + s) * 16, 0);
    return make_uint4(v[0], v[1], v[2], v[3]);
}

// Row z of each lane's SNP (voff = 16 * SNP index).
__device__ __forceinline__ uint4 ld_lane(__amdgpu_buffer_rsrc_t r, int np, int z, int voff)
{
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, z * np * 16, 0);
    return make_uint4(v[0], v[1], v[2], v[3]);
}

// acc_k += popcount(x_k & y_k & z_k) for four independent cells: one
// v_bitop3_b32 (3-input AND) and one accumulating v_bcnt_u32_b32 per cell,
// the four ANDs issued before the four popcounts.
__device__ __forceinline__ void and3_bcnt4(uint32_t &a0, uint32_t &a1, uint32_t &a2, uint32_t &a3, uint32_t x0,
                                           uint32_t y0, uint32_t z0, uint32_t x1, uint32_t y1, uint32_t z1,
                                           uint32_t x2, uint32_t y2, uint32_t z2, uint32_t x3, uint32_t y3,
                                           uint32_t z3)
{
    uint32_t t0, t1, t2, t3;
    asm("v_bitop3_b32 %0, %4, %5, %6 bitop3:0x80\n\t"
        "v_bitop3_b32 %1, %7, %8, %9 bitop3:0x80\n\t"
        "v_bitop3_b32 %2, %10, %11, %12 bitop3:0x80\n\t"
        "v_bitop3_b32 %3, %13, %14, %15 bitop3:0x80"
        : "=&v"(t0), "=&v"(t1), "=&v"(t2), "=&v"(t3)
        : "v"(x0), "v"(y0), "v"(z0), "v"(x1), "v"(y1), "v"(z1), "v"(x2), "v"(y2), "v"(z2), "v"(x3), "v"(y3),
          "v"(z3));
    asm("v_bcnt_u32_b32 %0, %4, %0\n\t"
        "v_bcnt_u32_b32 %1, %5, %1\n\t"
        "v_bcnt_u32_b32 %2, %6, %2\n\t"
        "v_bcnt_u32_b32 %3, %7, %3"
        : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3)
        : "v"(t0), "v"(t1), "v"(t2), "v"(t3));
}

// acc_k += popcount(x_k & y_k), four independent cells, VGPR operands.
__device__ __forceinline__ void and2_bcnt4(uint32_t &a0, uint32_t &a1, uint32_t &a2, uint32_t &a3, uint32_t x0,
                                           uint32_t y0, uint32_t x1, uint32_t y1, uint32_t x2, uint32_t y2,
                                           uint32_t x3, uint32_t y3)
{
    uint32_t t0, t1, t2, t3;
    asm("v_and_b32 %0, %4, %5\n\t"
        "v_and_b32 %1, %6, %7\n\t"
        "v_and_b32 %2, %8, %9\n\t"
        "v_and_b32 %3, %10, %11"
        : "=&v"(t0), "=&v"(t1), "=&v"(t2), "=&v"(t3)
        : "v"(x0), "v"(y0), "v"(x1), "v"(y1), "v"(x2), "v"(y2), "v"(x3), "v"(y3));
    asm("v_bcnt_u32_b32 %0, %4, %0\n\t"
        "v_bcnt_u32_b32 %1, %5, %1\n\t"
        "v_bcnt_u32_b32 %2, %6, %2\n\t"
        "v_bcnt_u32_b32 %3, %7, %3"
        : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3)
        : "v"(t0), "v"(t1), "v"(t2), "v"(t3));
}

// Lexicographic unranking of an m-combination of [0, n) (wave-uniform).
// rank(c) = C(n,m) - 1 - sum_i C(n-1-c_i, m-i), so the complement is written
// greedily in the combinatorial number system.
__device__ inline void unrank(uint64_t r, int n, int m, const uint64_t *binom, int *out)
{
    const uint64_t stride = (uint64_t)n + 1;
    uint64_t x = binom[(uint64_t)m * stride + n] - 1 - r;
    int hi = n - 1;
    for (int i = 0; i < m; ++i) {
        const int j = m - i;
        int lo = 0, h = hi;
        while (lo < h) {
            const int mid = (lo + h + 1) >> 1;
            if (binom[(uint64_t)j * stride + mid] <= x) lo = mid;
            else h = mid - 1;
        }
        x -= binom[(uint64_t)j * stride + lo];
        out[i] = n - 1 - lo;
        hi = lo - 1;
    }
}

// Lexicographic rank of the k-combination (pre[0] < ... < pre[k-2] < c) of
// [0, n), the key of a candidate: rank = C(n,k) - 1 - sum_i C(n-1-c_i, k-i),
// whose last term is C(n-1-c, 1) = n-1-c, so the rank is a constant of the
// prefix plus c.  Needs the binomial table up to row k and C(n, k) < 2^64
// (checked on the host).  Ranks keep the lexicographic order of the
// combinations, so comparing keys compares combinations.
__device__ __forceinline__ uint64_t prefix_key(const uint64_t *binom, int n, const int *pre, int M)
{
    const uint64_t stride = (uint64_t)n + 1;
    uint64_t s = binom[(uint64_t)(M + 1) * stride + n] - (uint64_t)n;   // C(n,k) - 1 - (n - 1)
    for (int i = 0; i < M; ++i) s -= binom[(uint64_t)(M + 1 - i) * stride + (n - 1 - pre[i])];
    return s;
}

// Screening MI in FP32 from exact counts, in the conditional-entropy form
// MI = H(Y) + (1/N) sum_x [n_cx ln(n_cx/n_x) + n_tx ln(n_tx/n_x)], whose terms
// all have one sign (no cancellation inside the sum).  Error < 2e-6 nats for
// N < 2^24; only used to decide which combinations get the exact value.
// Counts as uint32_t or as FP32 holding the same integers (exact below 2^24).
// The sum over cells, in log2 units (the screen of a table split over
// several lanes adds their sums).
// Cells come from cell(x, fc, ft) (counts as FP32), x < CELLS.
template <int CELLS, class F>
__device__ __forceinline__ float screen_sum_fn(F &&cell)
{
    // hardware v_rcp_f32 / v_log_f32 (log2) on normal arguments: every ratio
    // is clamped to >= 2^-100, and a zero count multiplies its log by zero
    float s = 0.0f;
#pragma unroll
    for (int x = 0; x < CELLS; ++x) {
        float fc, ft;
        cell(x, fc, ft);
        const float r = __builtin_amdgcn_rcpf(fmaxf(fc + ft, 1.0f));
        const float lc = __builtin_amdgcn_logf(fmaxf(fc * r, 7.9e-31f));
        const float lt = __builtin_amdgcn_logf(fmaxf(ft * r, 7.9e-31f));
        s = fmaf(fc, lc, s);
        s = fmaf(ft, lt, s);
    }
    return s;
}

template <int CELLS, class T>
__device__ __forceinline__ float screen_sum(const T *a0, const T *a1)
{
    return screen_sum_fn<CELLS>([&](int x, float &fc, float &ft) __attribute__((always_inline)) {
        fc = (float)a0[x];
        ft = (float)a1[x];
    });
}

template <int CELLS, class T>
__device__ __forceinline__ float mi_screen(const T *a0, const T *a1, float hy, float inv_n)
{
    return fmaf(screen_sum<CELLS>(a0, a1), inv_n * 0.69314718f, hy);
}

// Upper bound on the MI in nats, one reciprocal per cell and no logarithm:
// MI = KL(P_XY || P_X P_Y) <= ln(1 + chi2/N) <= chi2/N (Jensen's inequality
// on E_P[P_XY / (P_X P_Y)]), and with two classes
//   chi2/N = N/(n_c n_t) sum_x d_x^2 / n_x,  d_x = n_cx - n_x n_c/N.
// FP32 error: below 2e-6 relative plus 5e-7 absolute (each |d_x| error is
// under 1.2e-7 (n_x n_c/N + |d_x|), and sum_x |d_x| <= 2 min(n_c, n_t)).
// Cells from cell(x, fc, fx): case count and case + control count, FP32.
template <int CELLS, class F>
__device__ __forceinline__ float mi_bound_fn(F &&cell, float frac_c, float chi_scale)
{
    float s = 0.0f;
#pragma unroll
    for (int x = 0; x < CELLS; ++x) {
        float fc, fx;
        cell(x, fc, fx);
        const float d = fmaf(-fx, frac_c, fc);
        s = fmaf(d * d, __builtin_amdgcn_rcpf(fmaxf(fx, 1.0f)), s);
    }
    return s * chi_scale;
}

template <int CELLS, class T>
__device__ __forceinline__ float mi_bound(const T *a0, const T *a1, float frac_c, float chi_scale)
{
    return mi_bound_fn<CELLS>(
        [&](int x, float &fc, float &fx) __attribute__((always_inline)) {
            fx = (float)(a0[x] + a1[x]);
            fc = (float)a0[x];
        },
        frac_c, chi_scale);
}

// The bound rejects a combination when even its bound, raised by far more
// than the bound's own FP32 error plus the screen's (2e-6), stays below the
// screening threshold: the screen would then reject it as well, so the set
// of combinations that reach the exact MI is unchanged.
__device__ __forceinline__ bool bound_passes(float bound, float tau, float delta)
{
    return fmaf(bound, 1.0001f, 4e-6f) >= tau - delta;
}

// p*ln(p) in FP64 for p > 0, else 0 (kept out of line: it is only reached for
// the few combinations that pass the FP32 screen).
__device__ __noinline__ double plogp(double p)
{
#pragma clang fp contract(off)
    return p > 0.0 ? p * log(p) : 0.0;
}

// Exact MI: H(X) + H(Y) - H(X,Y) in FP64 from the integer counts, summed in
// cell order with no fused multiply-add, so that the value depends only on
// the counts (not on tiling, batching or launch shape).
// The running sums continue over CELLS consecutive cells (a table split over
// lanes in cell order is one chain through them).
// Cells from cell(x, n0, n1) (integer counts).
template <int CELLS, class F>
__device__ __forceinline__ void mi_exact_part_fn(F &&cell, double n, double &hx, double &hall)
{
#pragma clang fp contract(off)
#pragma unroll
    for (int x = 0; x < CELLS; ++x) {
        uint32_t n0, n1;
        cell(x, n0, n1);
        const double pc = (double)n0 / n, pt = (double)n1 / n, pa = pc + pt;
        hall -= plogp(pc);
        hall -= plogp(pt);
        hx -= plogp(pa);
    }
}

template <int CELLS>
__device__ __forceinline__ void mi_exact_part(const uint32_t *a0, const uint32_t *a1, double n, double &hx,
                                              double &hall)
{
    mi_exact_part_fn<CELLS>(
        [&](int x, uint32_t &n0, uint32_t &n1) __attribute__((always_inline)) {
            n0 = a0[x];
            n1 = a1[x];
        },
        n, hx, hall);
}

template <int CELLS>
__device__ __forceinline__ double mi_exact(const uint32_t *a0, const uint32_t *a1, double n, double hy)
{
#pragma clang fp contract(off)
    double hx = 0.0, hall = 0.0;
    mi_exact_part<CELLS>(a0, a1, n, hx, hall);
    return hx + hy - hall;
}

__device__ __forceinline__ bool better(float ma, uint64_t ca, float mb, uint64_t cb)
{
    return ma > mb || (ma == mb && ca < cb);
}

// Wave-aggregated push of candidates (all 64 lanes must call).
__device__ __forceinline__ void push_candidates(bool want, float mi, uint64_t combo,
                                                const ScanParams &P, TopState *st)
{
    const uint64_t m = __ballot(want);
    if (m == 0) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&st->cand_count, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    if (want) {
        const uint32_t idx = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        if (idx < P.cand_cap) {
            Cand c;
            c.mi = mi;
            c.tag = P.tag;
            c.combo = combo;
            P.cand[idx] = c;
        } else {
            atomicOr(&st->overflow, 1u);
        }
    }
}

// Screening MI, then exact MI and the key test for the survivors (one lane's
// combination; the push is separate).
template <int CELLS, class T, class K>
__device__ __forceinline__ void screen_exact(const T *a0, const T *a1, K &&key, float tau, uint64_t tau_combo,
                                             const ScanParams &P, bool &want, float &mf, uint64_t &combo)
{
    const float ms = mi_screen<CELLS>(a0, a1, P.hy, P.inv_n);
    if (ms >= tau - P.delta) {
        if constexpr (std::is_same<T, uint32_t>::value) {
            mf = (float)mi_exact<CELLS>(a0, a1, P.n64, P.hy64);
        } else {
            uint32_t u0[CELLS], u1[CELLS];
#pragma unroll
            for (int x = 0; x < CELLS; ++x) u0[x] = (uint32_t)a0[x], u1[x] = (uint32_t)a1[x];
            mf = (float)mi_exact<CELLS>(u0, u1, P.n64, P.hy64);
        }
        combo = key();
        want = better(mf, combo, tau, tau_combo);
    }
}

// Whether a valid combination passes the bound (always, without it).
template <int CELLS, class T>
__device__ __forceinline__ bool bound_ok(const T *a0, const T *a1, float tau, const ScanParams &P)
{
    return !P.bounded || bound_passes(mi_bound<CELLS>(a0, a1, P.frac_c, P.chi_scale), tau, P.delta);
}

// Epilogue shared by every scan kernel: screening MI, exact MI for the
// survivors, key test against the current N-th best, push.  key() gives the
// combination's key; it is only called for the few combinations that reach
// the exact MI.
template <int CELLS, class T, class K>
__device__ __forceinline__ void score_and_push(bool valid, const T *a0, const T *a1, K &&key, float tau,
                                               uint64_t tau_combo, const ScanParams &P, TopState *st)
{
    bool want = false;
    float mf = 0.0f;
    uint64_t combo = ~0ull;
    if (valid && bound_ok<CELLS>(a0, a1, tau, P)) screen_exact<CELLS>(a0, a1, key, tau, tau_combo, P, want, mf, combo);
    push_candidates(want, mf, combo, P, st);
}

// Uniform lookup of prefix j < j1: rank -> SNP indices, or an empty marker.
template <int M>
__device__ __forceinline__ bool prefix_at(const ScanParams &P, int n, int64_t j, int64_t j1, int *pre)
{
    if (j >= j1) return false;
    const uint64_t rank = (uint64_t)P.offset + (uint64_t)j * (uint64_t)P.step;
    if (rank >= P.total_prefixes) return false;
    unrank(rank, n, M, P.binom, pre);
    return true;
}

// Either the normal epilogue, or (threshold sampling pass) one count in the
// workgroup's LDS histogram of screening MI values.
template <bool HIST, int CELLS, class K>
__device__ __forceinline__ void score_or_sample(bool valid, const uint32_t *a0, const uint32_t *a1, K &&key,
                                                float tau, uint64_t tau_combo, const ScanParams &P, TopState *st,
                                                uint32_t *lds_hist)
{
    if constexpr (HIST) {
        if (valid) atomicAdd(&lds_hist[hist_bin(mi_screen<CELLS>(a0, a1, P.hy, P.inv_n))], 1u);
    } else {
        score_and_push<CELLS>(valid, a0, a1, key, tau, tau_combo, P, st);
    }
}

template <bool HIST>
__device__ __forceinline__ void hist_begin(uint32_t *lds_hist)
{
    if constexpr (HIST) {
        for (int b = threadIdx.x; b < kHistBins; b += blockDim.x) lds_hist[b] = 0;
        __syncthreads();
    }
}

template <bool HIST>
__device__ __forceinline__ void hist_end(uint32_t *lds_hist, uint32_t *hist)
{
    if constexpr (HIST) {
        __syncthreads();
        for (int b = threadIdx.x; b < kHistBins; b += blockDim.x)
            if (lds_hist[b]) atomicAdd(&hist[b], lds_hist[b]);
    }
}

// Starting threshold from the sampling histogram: the lower edge of the bin
// holding the N-th best sampled screening value, less the screening margin.
// The sampled combinations are scanned again afterwards, and at least N of
// them have an exact value above it, so the final N-th best is never below.
// One workgroup of kTauThreads: thread t sums bins [kTauPer t, kTauPer (t +
// 1)), a suffix scan over the threads gives the count at or above each
// thread's range, and the thread whose range holds the crossing walks its
// own bins (the histogram is read once, in parallel).
constexpr int kTauThreads = 256, kTauPer = kHistBins / kTauThreads;
static_assert(kHistBins % kTauThreads == 0, "histogram split");
__global__ void __launch_bounds__(kTauThreads) k_tau_from_hist(const uint32_t *hist, uint32_t N, float delta,
                                                                TopState *st)
{
    __shared__ uint64_t part[kTauThreads];
    const int t = threadIdx.x;
    uint32_t h[kTauPer];
    uint64_t mine = 0;
#pragma unroll
    for (int i = 0; i < kTauPer; ++i) {
        h[i] = hist[t * kTauPer + i];
        mine += h[i];
    }
    part[t] = mine;
    __syncthreads();
    // inclusive suffix sums: part[t] = sum over threads t' >= t
    for (int d = 1; d < kTauThreads; d <<= 1) {
        const uint64_t add = t + d < kTauThreads ? part[t + d] : 0;
        __syncthreads();
        part[t] += add;
        __syncthreads();
    }
    const uint64_t above = t + 1 < kTauThreads ? part[t + 1] : 0;   // bins past this thread's range
    if (N == 0 || above >= N || above + mine < N) return;
    uint64_t cum = above;
    for (int i = kTauPer - 1; i >= 0; --i) {
        cum += h[i];
        if (cum >= N) {
            union {
                uint32_t u;
                float f;
            } x;
            x.u = (uint32_t)(t * kTauPer + i) << 20;
            st->tau = x.f - delta;
            st->tau_combo = ~0ull;
            return;
        }
    }
}

// ---------------------------------------------------------------------------
// Direct scan: all 3^k cells by AND + popcount (the reference's algorithm).
// ---------------------------------------------------------------------------
// One 32-bit word of every cell: cell (g_0 .. g_{k-1}) gets
// popcount(row g_0 of the first SNP & ... & row g_{k-1} of the lane's SNP).
// Orders 2 and 3 take one v_and_b32 / v_bitop3_b32 per cell; order 4 first
// forms the 9 pair words of the first two SNPs (9 VALU per word for 81 cells).
template <int K>
__device__ __forceinline__ void count_word(uint32_t *acc, const uint32_t (&p)[K - 1][3], const uint32_t (&x)[3])
{
    constexpr int CELLS = Pow3<K>::v;
    uint32_t o0[CELLS], o1[CELLS], o2[CELLS];
    if constexpr (K == 2) {
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int z = 0; z < 3; ++z) o0[a * 3 + z] = p[0][a], o1[a * 3 + z] = x[z];
    } else if constexpr (K == 3) {
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b)
#pragma unroll
                for (int z = 0; z < 3; ++z)
                    o0[(a * 3 + b) * 3 + z] = p[0][a], o1[(a * 3 + b) * 3 + z] = p[1][b],
                                         o2[(a * 3 + b) * 3 + z] = x[z];
    } else {
        uint32_t ab[9];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) ab[a * 3 + b] = p[0][a] & p[1][b];
#pragma unroll
        for (int r = 0; r < 9; ++r)
#pragma unroll
            for (int e = 0; e < 3; ++e)
#pragma unroll
                for (int z = 0; z < 3; ++z)
                    o0[(r * 3 + e) * 3 + z] = ab[r], o1[(r * 3 + e) * 3 + z] = p[2][e], o2[(r * 3 + e) * 3 + z] = x[z];
    }
    constexpr int FULL = CELLS / 4 * 4;
#pragma unroll
    for (int i = 0; i < FULL; i += 4) {
        if constexpr (K == 2)
            and2_bcnt4(acc[i], acc[i + 1], acc[i + 2], acc[i + 3], o0[i], o1[i], o0[i + 1], o1[i + 1], o0[i + 2],
                       o1[i + 2], o0[i + 3], o1[i + 3]);
        else
            and3_bcnt4(acc[i], acc[i + 1], acc[i + 2], acc[i + 3], o0[i], o1[i], o2[i], o0[i + 1], o1[i + 1],
                       o2[i + 1], o0[i + 2], o1[i + 2], o2[i + 2], o0[i + 3], o1[i + 3], o2[i + 3]);
    }
#pragma unroll
    for (int i = FULL; i < CELLS; ++i) {
        if constexpr (K == 2) acc[i] += __builtin_popcount(o0[i] & o1[i]);
        else acc[i] += __builtin_popcount(o0[i] & o1[i] & o2[i]);
    }
}

// All 3^k cells of one class for the lanes' SNPs, both the lanes' rows and
// the prefix rows of the next quad in flight while the current quad counts.
template <int K>
__device__ __forceinline__ void accumulate_direct(const uint4 *__restrict__ base, int np, int nq,
                                                  const int *pre, int voff, uint32_t *acc)
{
    constexpr int M = K - 1, CELLS = Pow3<K>::v;
#pragma unroll
    for (int i = 0; i < CELLS; ++i) acc[i] = 0;
    if (nq <= 0) return;
    const uint4 *row = base;
    __amdgpu_buffer_rsrc_t r = quad_rsrc(row, np);
    uint4 cz[3], pz[M][3];
#pragma unroll
    for (int z = 0; z < 3; ++z) cz[z] = ld_lane(r, np, z, voff);
#pragma unroll
    for (int i = 0; i < M; ++i)
#pragma unroll
        for (int z = 0; z < 3; ++z) pz[i][z] = ld_uniform(r, np, z, pre[i]);
    for (int q = 0; q < nq; ++q) {
        const uint4 *nrow = q + 1 < nq ? row + (size_t)3 * np : row;
        const __amdgpu_buffer_rsrc_t nr = quad_rsrc(nrow, np);
        uint4 ncz[3], npz[M][3];
#pragma unroll
        for (int z = 0; z < 3; ++z) ncz[z] = ld_lane(nr, np, z, voff);
#pragma unroll
        for (int i = 0; i < M; ++i)
#pragma unroll
            for (int z = 0; z < 3; ++z) npz[i][z] = ld_uniform(nr, np, z, pre[i]);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            uint32_t p[M][3], x[3];
#pragma unroll
            for (int i = 0; i < M; ++i)
#pragma unroll
                for (int z = 0; z < 3; ++z) p[i][z] = comp(pz[i][z], w);
#pragma unroll
            for (int z = 0; z < 3; ++z) x[z] = comp(cz[z], w);
            count_word<K>(acc, p, x);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int z = 0; z < 3; ++z) cz[z] = ncz[z];
#pragma unroll
        for (int i = 0; i < M; ++i)
#pragma unroll
            for (int z = 0; z < 3; ++z) pz[i][z] = npz[i][z];
        r = nr;
        row = nrow;
    }
}

// One wavefront's unit of a direct scan: prefix j of the distribution against
// the tiles t_hi, t_hi - 1, ..., down to max(t_lo, first tile past the
// prefix).  Whole-prefix units pass t_hi = last tile, t_lo = 0.
template <int K, bool HIST>
__device__ __forceinline__ void direct_unit(const DevData &D, const ScanParams &P, TopState *st, int64_t j,
                                            int64_t j1, int t_hi, int t_lo, uint32_t *lds_hist)
{
    constexpr int M = K - 1, CELLS = Pow3<K>::v;
    const int lane = threadIdx.x & 63;
    int pre[M];
    if (!prefix_at<M>(P, D.n, j, j1, pre)) return;
    const uint64_t pre_key = prefix_key(P.binom, D.n, pre, M);
    const int last = pre[M - 1];
    const float tau = st->tau;
    const uint64_t tau_combo = st->tau_combo;
    const uint4 *b0 = D.bits + D.class_off[0];
    const uint4 *b1 = D.bits + D.class_off[1];
    // tiles are walked from the last one down, so that wavefronts resident at
    // the same time read the same tile at about the same time (shared in L2)
    const int t_end = max(t_lo, (last + 1) >> 6);
    for (int t = min(t_hi, (D.n_pad >> 6) - 1); t >= t_end; --t) {
        const int c = (t << 6) | lane;
        uint32_t a0[CELLS], a1[CELLS];
        accumulate_direct<K>(b0, D.n_pad, D.quads[0], pre, c * 16, a0);
        accumulate_direct<K>(b1, D.n_pad, D.quads[1], pre, c * 16, a1);
        score_or_sample<HIST, CELLS>(c > last && c < D.n, a0, a1, [&] { return pre_key + (uint64_t)c; }, tau,
                                     tau_combo, P, st, lds_hist);
    }
}

template <int K, bool HIST>
__global__ void __launch_bounds__(kBlock)
k_scan_direct(const DevData D, const ScanParams P, TopState *st)
{
    __shared__ uint32_t lds_hist[HIST ? kHistBins : 1];
    hist_begin<HIST>(lds_hist);
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    direct_unit<K, HIST>(D, P, st, P.j0 + (int64_t)blockIdx.x * (kBlock / kWave) + wid, P.j1, INT32_MAX, 0,
                         lds_hist);
    hist_end<HIST>(lds_hist, P.hist);  // every wave of the block reaches the barrier
}

// ---------------------------------------------------------------------------
// Orders 7 to 9 (FIU_ALGO_DIRECT: all cells, any dataset).
// ---------------------------------------------------------------------------
// A workgroup takes one prefix (s_0 < ... < s_{k-2}) of the distribution and
// runs through its last SNPs c.  Every one of the 3^k cells is counted as the
// reference counts it: the AND of one genotype row per SNP, then popcount
// (gt_combine.cpp:3-27, gt_popcnt.cpp:10-18).  Thread t < 243 owns the
// genotypes of the first five SNPs (t = sum g_q 3^(4-q)); at orders 8 and 9 a
// loop takes those of the middle SNPs (the sixth, and the seventh); per
// 128-bit quad the thread counts the 9 cells of the last two SNPs.  Cell x = sum g_q 3^(k-1-q) (ContingencyTable.h:67-90); the
// counts of both classes go to LDS.  The MI is then summed in FP64 over the
// cells in any order (all threads, then a reduction).  Its difference from
// the cell-order value is below 3^k x 2^-52 x sum |p ln p| < 1e-10 nats, so
// a combination whose value plus 1e-9 rounds below the current threshold
// would be rejected by its exact value as well.  The others get the exact
// value in cell order (mi_exact's chain: wave 0 computes the terms of 64
// cells at a time and every lane runs the chain over them).
template <int K, bool HIST>
__global__ void __launch_bounds__(kBlock) k_scan_any(const DevData D, const ScanParams P, TopState *st)
{
    static_assert(K >= 7 && K <= 9, "orders 7 to 9");   // order 9: 157.5 KB of LDS
    constexpr int M = K - 1, CELLS = Pow3<K>::v, PASSES = Pow3<K - 7>::v;
    __shared__ uint32_t cnt[2][CELLS];
    __shared__ double red[2][kBlock / kWave];
    __shared__ int pass_sh;
    const int64_t j = P.j0 + blockIdx.x;
    int s[K];
    if (!prefix_at<M>(P, D.n, j, P.j1, s)) return;   // uniform over the workgroup
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const float tau = st->tau;
    const uint64_t tau_combo = st->tau_combo;
    const int np = D.n_pad;
    int g5[5];
    {
        int t = tid < 243 ? tid : 0;
#pragma unroll
        for (int q = 4; q >= 0; --q) g5[q] = t % 3, t /= 3;
    }
    const uint64_t pre_key = prefix_key(P.binom, D.n, s, M);
    for (int c = s[M - 1] + 1; c < D.n; ++c) {
        s[M] = c;
#pragma unroll
        for (int cls = 0; cls < 2; ++cls) {
            uint32_t n9[PASSES][9];
#pragma unroll
            for (int pi = 0; pi < PASSES; ++pi)
#pragma unroll
                for (int z = 0; z < 9; ++z) n9[pi][z] = 0;
            const uint4 *base = D.bits + D.class_off[cls];
            for (int q = 0; q < D.quads[cls]; ++q) {
                // row z of SNP x in quad q: rq[z * n_pad + x]; the prefix row
                // of the first five SNPs once, then every middle genotype
                const uint4 *rq = base + (uint64_t)q * 3 * np;
                uint4 pr = rq[g5[0] * np + s[0]];
#pragma unroll
                for (int i = 1; i < 5; ++i) pr = pr & rq[g5[i] * np + s[i]];
                uint4 ra[3], rb[3];
#pragma unroll
                for (int z = 0; z < 3; ++z) ra[z] = rq[z * np + s[K - 2]], rb[z] = rq[z * np + s[K - 1]];
#pragma unroll
                for (int pi = 0; pi < PASSES; ++pi) {
                    uint4 pm = pr;
                    if constexpr (K >= 8) pm = pm & rq[(K == 8 ? pi : pi / 3) * np + s[5]];
                    if constexpr (K == 9) pm = pm & rq[(pi % 3) * np + s[6]];
#pragma unroll
                    for (int z1 = 0; z1 < 3; ++z1) {
                        const uint4 pz = pm & ra[z1];
#pragma unroll
                        for (int z2 = 0; z2 < 3; ++z2) {
                            const uint4 v = pz & rb[z2];
                            n9[pi][3 * z1 + z2] += __builtin_popcount(v.x) + __builtin_popcount(v.y) +
                                                   __builtin_popcount(v.z) + __builtin_popcount(v.w);
                        }
                    }
                }
            }
            if (tid < 243) {
#pragma unroll
                for (int pi = 0; pi < PASSES; ++pi)
#pragma unroll
                    for (int z = 0; z < 9; ++z) cnt[cls][(tid * PASSES + pi) * 9 + z] = n9[pi][z];
            }
        }
        __syncthreads();
        // the MI in any order (the screen)
        double hx = 0.0, hall = 0.0;
        for (int x = tid; x < CELLS; x += kBlock) {
            const double pc = (double)cnt[0][x] / P.n64, pt = (double)cnt[1][x] / P.n64;
            hall += plogp(pc) + plogp(pt);
            hx += plogp(pc + pt);
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            hx += __shfl_xor(hx, off);
            hall += __shfl_xor(hall, off);
        }
        if (lane == 0) red[0][wv] = hx, red[1][wv] = hall;
        const uint64_t combo = pre_key + (uint64_t)c;
        __syncthreads();
        if (tid == 0) {
            double sx = 0.0, sa = 0.0;
#pragma unroll
            for (int w = 0; w < kBlock / kWave; ++w) sx += red[0][w], sa += red[1][w];
            const double v = P.hy64 - sx + sa;
            if constexpr (HIST) atomicAdd(&P.hist[hist_bin((float)v)], 1u);
            pass_sh = !HIST && (float)(v + 1e-9) >= tau;
        }
        __syncthreads();
        if (pass_sh && wv == 0) {
#pragma clang fp contract(off)
            double ex = 0.0, ea = 0.0;
            for (int x0 = 0; x0 < CELLS; x0 += kWave) {
                const int x = x0 + lane;
                double tc = 0.0, tt = 0.0, ta = 0.0;
                if (x < CELLS) {
                    const double pc = (double)cnt[0][x] / P.n64, pt = (double)cnt[1][x] / P.n64, pa = pc + pt;
                    tc = plogp(pc), tt = plogp(pt), ta = plogp(pa);
                }
                const int last = CELLS - x0 < kWave ? CELLS - x0 : kWave;
                for (int l = 0; l < last; ++l) {
                    const double a = __shfl(tc, l), b = __shfl(tt, l), e = __shfl(ta, l);
                    ea -= a;
                    ea -= b;
                    ex -= e;
                }
            }
            const float mf = (float)(ex + P.hy64 - ea);
            push_candidates(lane == 0 && better(mf, combo, tau, tau_combo), mf, combo, P, st);
        }
        __syncthreads();   // cnt and pass_sh are rewritten for the next c
    }
}

// ---------------------------------------------------------------------------
// Orders 10 to 12 (FIU_ALGO_DIRECT: all cells, any dataset).
// ---------------------------------------------------------------------------
// The same counting as k_scan_any, with the 3^k cells of both classes in
// global memory (472 KB per combination at order 10, 4.25 MB at order 12):
// the launch is persistent, workgroup w owns slot w of `cells` (2 x 3^k
// counters) and takes the prefixes j0 + w, j0 + w + gridDim.x, ... of the
// chunk, so that no two resident workgroups share a slot.  Thread t < 243
// owns the genotypes of the first five SNPs; the 3^(k-7) genotypes of the
// middle SNPs (the sixth to the (k-2)-th) are taken 9 at a time, each block
// re-reading the quads (they stay in L2); per 128-bit quad the thread counts
// the 9 cells of the last two SNPs.  Cell x = sum g_q 3^(k-1-q) = (t 3^(k-7) +
// p) 9 + z, p the middle genotypes' index, z those of the last two.  Screen
// and exact MI as in k_scan_any, reading the counts from the slot.
template <int K, bool HIST>
__global__ void __launch_bounds__(kBlock) k_scan_big(const DevData D, const ScanParams P, TopState *st,
                                                     uint32_t *cells)
{
    static_assert(K >= 10 && K <= 12, "orders 10 to 12");
    constexpr int M = K - 1, CELLS = Pow3<K>::v, PASSES = Pow3<K - 7>::v, PB = 9, MID = K - 7;
    __shared__ double red[2][kBlock / kWave];
    __shared__ int pass_sh;
    uint32_t *cnt[2] = {cells + (uint64_t)blockIdx.x * 2 * CELLS, cells + (uint64_t)blockIdx.x * 2 * CELLS + CELLS};
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int np = D.n_pad;
    int g5[5];
    {
        int t = tid < 243 ? tid : 0;
#pragma unroll
        for (int q = 4; q >= 0; --q) g5[q] = t % 3, t /= 3;
    }
    for (int64_t j = P.j0 + blockIdx.x; j < P.j1; j += gridDim.x) {
        int s[K];
        if (!prefix_at<M>(P, D.n, j, P.j1, s)) break;   // uniform; later j lie past the last prefix too
        const float tau = st->tau;
        const uint64_t tau_combo = st->tau_combo;
        const uint64_t pre_key = prefix_key(P.binom, D.n, s, M);
        for (int c = s[M - 1] + 1; c < D.n; ++c) {
            s[M] = c;
            for (int cls = 0; cls < 2; ++cls) {
                const uint4 *base = D.bits + D.class_off[cls];
                for (int p0 = 0; p0 < PASSES; p0 += PB) {
                    uint32_t n9[PB][9];
#pragma unroll
                    for (int pi = 0; pi < PB; ++pi)
#pragma unroll
                        for (int z = 0; z < 9; ++z) n9[pi][z] = 0;
                    for (int q = 0; q < D.quads[cls]; ++q) {
                        const uint4 *rq = base + (uint64_t)q * 3 * np;
                        uint4 pr = rq[g5[0] * np + s[0]];
#pragma unroll
                        for (int i = 1; i < 5; ++i) pr = pr & rq[g5[i] * np + s[i]];
                        uint4 ra[3], rb[3];
#pragma unroll
                        for (int z = 0; z < 3; ++z) ra[z] = rq[z * np + s[K - 2]], rb[z] = rq[z * np + s[K - 1]];
#pragma unroll
                        for (int pi = 0; pi < PB; ++pi) {
                            // middle genotypes of pass p0 + pi, the first middle SNP most significant
                            uint4 pm = pr;
                            int p = p0 + pi;
#pragma unroll
                            for (int i = MID - 1; i >= 0; --i) {
                                pm = pm & rq[(p % 3) * np + s[5 + i]];
                                p /= 3;
                            }
#pragma unroll
                            for (int z1 = 0; z1 < 3; ++z1) {
                                const uint4 pz = pm & ra[z1];
#pragma unroll
                                for (int z2 = 0; z2 < 3; ++z2) {
                                    const uint4 v = pz & rb[z2];
                                    n9[pi][3 * z1 + z2] += __builtin_popcount(v.x) + __builtin_popcount(v.y) +
                                                           __builtin_popcount(v.z) + __builtin_popcount(v.w);
                                }
                            }
                        }
                    }
                    if (tid < 243) {
#pragma unroll
                        for (int pi = 0; pi < PB; ++pi)
#pragma unroll
                            for (int z = 0; z < 9; ++z) cnt[cls][(tid * PASSES + p0 + pi) * 9 + z] = n9[pi][z];
                    }
                }
            }
            __threadfence_block();
            __syncthreads();
            // the MI in any order (the screen), as in k_scan_any
            double hx = 0.0, hall = 0.0;
            for (int x = tid; x < CELLS; x += kBlock) {
                const double pc = (double)cnt[0][x] / P.n64, pt = (double)cnt[1][x] / P.n64;
                hall += plogp(pc) + plogp(pt);
                hx += plogp(pc + pt);
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                hx += __shfl_xor(hx, off);
                hall += __shfl_xor(hall, off);
            }
            if (lane == 0) red[0][wv] = hx, red[1][wv] = hall;
            const uint64_t combo = pre_key + (uint64_t)c;
            __syncthreads();
            if (tid == 0) {
                double sx = 0.0, sa = 0.0;
#pragma unroll
                for (int w = 0; w < kBlock / kWave; ++w) sx += red[0][w], sa += red[1][w];
                const double v = P.hy64 - sx + sa;
                if constexpr (HIST) atomicAdd(&P.hist[hist_bin((float)v)], 1u);
                pass_sh = !HIST && (float)(v + 1e-9) >= tau;
            }
            __syncthreads();
            if (pass_sh && wv == 0) {
#pragma clang fp contract(off)
                double ex = 0.0, ea = 0.0;
                for (int x0 = 0; x0 < CELLS; x0 += kWave) {
                    const int x = x0 + lane;
                    double tc = 0.0, tt = 0.0, ta = 0.0;
                    if (x < CELLS) {
                        const double pc = (double)cnt[0][x] / P.n64, pt = (double)cnt[1][x] / P.n64, pa = pc + pt;
                        tc = plogp(pc), tt = plogp(pt), ta = plogp(pa);
                    }
                    const int last = CELLS - x0 < kWave ? CELLS - x0 : kWave;
                    for (int l = 0; l < last; ++l) {
                        const double a = __shfl(tc, l), b = __shfl(tt, l), e = __shfl(ta, l);
                        ea -= a;
                        ea -= b;
                        ex -= e;
                    }
                }
                const float mf = (float)(ex + P.hy64 - ea);
                push_candidates(lane == 0 && better(mf, combo, tau, tau_combo), mf, combo, P, st);
            }
            __syncthreads();   // the slot and pass_sh are rewritten for the next c
        }
    }
}

// ---------------------------------------------------------------------------
// Orders 4 to 6, marginal in the last SNP (FIU_ALGO_MARGINAL).
// ---------------------------------------------------------------------------
// A wavefront takes one prefix (s_0 < ... < s_{k-2}) of the distribution and
// the last SNPs c after it.  The prefix's 3^(k-1) genotype rows (ANDs of one
// row per prefix SNP, as GenotypeTable::combine chains them,
// gt_combine.cpp:3-27) are split over L = 3^(k-3) lanes per c by the
// genotypes of the first k-3 prefix SNPs (the lane's part p); the 9 rows of
// the last two prefix SNPs are wave-uniform scalars built per 32-bit word
// from scalar loads.  A lane ANDs each of them with its part's row and with
// c's rows 0 and 1, and counts (AND + popcount, gt_popcnt.cpp:10-18).
// Genotype 2 of c follows from the prefix row's own count, taken once per
// prefix: c's rows partition the samples that every SNP covers (checked at
// upload).  Cells counted per word: 2 * 3^(k-1) instead of 3^k.  Cell
// x = 3 g + z with g = 9 p + r, first SNP most significant
// (ContingencyTable.h:67-90): the L lanes of a c hold consecutive cell
// ranges, so the bound and the screen are sums over them and the exact MI is
// one chain through them in cell order.  Lanes: order 4, 21 c x 3 parts;
// order 5, 7 c x 9 parts (lane 63 idles); order 6, 2 c x 27 parts (ten
// lanes idle).
template <int K> struct Km {
    static_assert(K >= 4 && K <= 6, "orders 4 to 6");
    static constexpr int LS = K - 3;            // prefix SNPs split over lanes
    static constexpr int L = Pow3<K - 3>::v;     // lanes per c
    static constexpr int CPW = 64 / L;           // c per wavefront
    static constexpr int NR = 9;                 // uniform rows (last two prefix SNPs)
    static constexpr int LC = 3 * NR;            // cells per lane
};

// One class: acc[r][z] = sum over words of popcount(row r & f & c_z), z < Z
// (Z = 3: c's three rows; Z = 2: c's rows 0 and 1; Z = 1: no c, the prefix
// rows' own counts).
// f = AND of rows digit_i(p) of the lane-split SNPs pre[i] (lane offsets
// voff_f[i]); row r = AND of rows r / 3 and r % 3 of the last two prefix
// SNPs (scalar loads: their indices are wave-uniform); c at voff_c.
template <int K, int Z>
__device__ __forceinline__ void km_count(const uint4 *__restrict__ base, int np, int nq, const int (&pre)[K - 1],
                                         const int (&voff_f)[Km<K>::LS], int voff_c, uint32_t (&acc)[9][Z])
{
    constexpr int LS = Km<K>::LS, NR = Km<K>::NR;
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int z = 0; z < Z; ++z) acc[r][z] = 0;
    if (nq <= 0) return;
    const int u0 = pre[K - 3], u1 = pre[K - 2];
    const uint4 *row = base;
    __amdgpu_buffer_rsrc_t rs = quad_rsrc(row, np);
    uint4 f[LS], cz[Z], pz[2][3];
    auto load = [&](const uint4 *rw, __amdgpu_buffer_rsrc_t r, uint4(&lf)[LS], uint4(&lc)[Z], uint4(&lp)[2][3])
        __attribute__((always_inline)) {
#pragma unroll
        for (int i = 0; i < LS; ++i) lf[i] = ld_lane(r, np, 0, voff_f[i]);
        if constexpr (Z >= 2) {
#pragma unroll
            for (int z = 0; z < Z; ++z) lc[z] = ld_lane(r, np, z, voff_c);
        }
#pragma unroll
        for (int z = 0; z < 3; ++z) {
            lp[0][z] = rw[z * np + u0];
            lp[1][z] = rw[z * np + u1];
        }
    };
    load(row, rs, f, cz, pz);
    for (int q = 0; q < nq; ++q) {
        const uint4 *nrow = q + 1 < nq ? row + (size_t)3 * np : row;
        const __amdgpu_buffer_rsrc_t nr = quad_rsrc(nrow, np);
        uint4 nf[LS], ncz[Z], npz[2][3];
        load(nrow, nr, nf, ncz, npz);
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            uint32_t rows[NR];
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) rows[a * 3 + b] = comp(pz[0][a], w) & comp(pz[1][b], w);
            uint32_t fw = comp(f[0], w);
#pragma unroll
            for (int i = 1; i < LS; ++i) fw &= comp(f[i], w);
            if constexpr (Z >= 2) {
                uint32_t x[Z];
#pragma unroll
                for (int z = 0; z < Z; ++z) x[z] = comp(cz[z], w) & fw;
#pragma unroll
                for (int r = 0; r < NR; ++r)
#pragma unroll
                    for (int z = 0; z < Z; ++z) acc[r][z] += __builtin_popcount(rows[r] & x[z]);
            } else {
#pragma unroll
                for (int r = 0; r < NR; ++r) acc[r][0] += __builtin_popcount(rows[r] & fw);
            }
        }
#pragma unroll
        for (int i = 0; i < LS; ++i) f[i] = nf[i];
#pragma unroll
        for (int z = 0; z < Z; ++z) cz[z] = ncz[z];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int z = 0; z < 3; ++z) pz[a][z] = npz[a][z];
        rs = nr;
        row = nrow;
    }
}

// Last SNPs c in tiles t_lo .. t_hi (64 SNPs each) only.  ALL: the three
// rows of c are counted (every one of the 3^k cells by AND + popcount, the
// reference's algorithm), which needs no coverage property of the dataset;
// otherwise genotype 2 of c follows from the prefix rows' own counts.
template <int K, bool HIST, bool ALL = false>
__device__ __forceinline__ void km_unit(const DevData &D, const ScanParams &P, TopState *st, int64_t j, int64_t j1,
                                        int t_hi, int t_lo, uint32_t *lds_hist)
{
    constexpr int M = K - 1, LS = Km<K>::LS, L = Km<K>::L, CPW = Km<K>::CPW, NR = Km<K>::NR, LC = Km<K>::LC;
    int pre[M];
    if (!prefix_at<M>(P, D.n, j, j1, pre)) return;
#pragma unroll
    for (int i = 0; i < M; ++i) pre[i] = __builtin_amdgcn_readfirstlane(pre[i]);
    const uint64_t pre_key = prefix_key(P.binom, D.n, pre, M);
    const int last = pre[M - 1], n = D.n, np = D.n_pad;
    const int c_begin = max(last + 1, 64 * t_lo), c_end = min(n, 64 * min(t_hi, (np >> 6) - 1) + 64);
    if (c_begin >= c_end) return;
    const int lane = threadIdx.x & 63, part = lane % L, first_lane = lane - part;
    const bool lane_ok = lane < CPW * L;
    int voff_f[LS];
#pragma unroll
    for (int i = 0; i < LS; ++i) {
        int digit = part;
#pragma unroll
        for (int e = i + 1; e < LS; ++e) digit /= 3;
        voff_f[i] = 16 * ((digit % 3) * np + pre[i]);
    }
    const uint4 *b0 = D.bits + D.class_off[0];
    const uint4 *b1 = D.bits + D.class_off[1];
    // the prefix rows' own counts (genotype "any" of c)
    constexpr int Z = ALL ? 3 : 2;
    uint32_t pr0[NR][1], pr1[NR][1];
    if constexpr (!ALL) {
        km_count<K, 1>(b0, np, D.quads[0], pre, voff_f, 0, pr0);
        km_count<K, 1>(b1, np, D.quads[1], pre, voff_f, 0, pr1);
    }
    const float tau = st->tau;
    const uint64_t tau_combo = st->tau_combo;
    // sums over the L lanes of this c, in the same order in each
    auto sum_parts = [&](float v) __attribute__((always_inline)) {
        float t = __shfl(v, first_lane);
#pragma unroll
        for (int i = 1; i < L; ++i) t += __shfl(v, first_lane + i);
        return t;
    };
    for (int c0 = c_begin; c0 < c_end; c0 += CPW) {
        const int c = c0 + lane / L;
        const bool valid = lane_ok && c < c_end;
        const int voff_c = 16 * min(c, np - 1);
        uint32_t k0[NR][Z], k1[NR][Z];
        km_count<K, Z>(b0, np, D.quads[0], pre, voff_f, voff_c, k0);
        km_count<K, Z>(b1, np, D.quads[1], pre, voff_f, voff_c, k1);
        // this lane's cell x = 3 r + z (z = 2, marginal form: the prefix
        // row's count minus genotypes 0 and 1), computed where it is used
        auto cnt = [&](int x, uint32_t &n0, uint32_t &n1) __attribute__((always_inline)) {
            const int r = x / 3, z = x % 3;
            if constexpr (ALL) {
                n0 = k0[r][z];
                n1 = k1[r][z];
            } else {
                n0 = z < 2 ? k0[r][z] : pr0[r][0] - k0[r][0] - k0[r][1];
                n1 = z < 2 ? k1[r][z] : pr1[r][0] - k1[r][0] - k1[r][1];
            }
        };
        auto cnt_screen = [&](int x, float &fc, float &ft) __attribute__((always_inline)) {
            uint32_t n0, n1;
            cnt(x, n0, n1);
            fc = (float)n0;
            ft = (float)n1;
        };
        auto cnt_bound = [&](int x, float &fc, float &fx) __attribute__((always_inline)) {
            uint32_t n0, n1;
            cnt(x, n0, n1);
            fc = (float)n0;
            fx = (float)(n0 + n1);
        };
        if constexpr (HIST) {
            const float ms = fmaf(sum_parts(screen_sum_fn<LC>(cnt_screen)), P.inv_n * 0.69314718f, P.hy);
            if (valid && part == 0) atomicAdd(&lds_hist[hist_bin(ms)], 1u);
        } else {
            bool pass = valid;
            if (P.bounded)
                pass = pass && bound_passes(sum_parts(mi_bound_fn<LC>(cnt_bound, P.frac_c, P.chi_scale)), tau, P.delta);
            float ms = 0.0f;
            if (__ballot(pass)) ms = fmaf(sum_parts(screen_sum_fn<LC>(cnt_screen)), P.inv_n * 0.69314718f, P.hy);
            const bool exact = pass && ms >= tau - P.delta;
            bool want = false;
            float mf = 0.0f;
            if (__ballot(exact)) {
                // one FP64 chain through the L lanes, in cell order
                double hx = 0.0, hall = 0.0;
                for (int p = 0; p < L; ++p) {
                    if (exact && part == p) mi_exact_part_fn<LC>(cnt, P.n64, hx, hall);
                    const double nhx = __shfl(hx, lane > 0 ? lane - 1 : 0);
                    const double nhall = __shfl(hall, lane > 0 ? lane - 1 : 0);
                    if (part == p + 1) {
                        hx = nhx;
                        hall = nhall;
                    }
                }
                if (exact && part == L - 1) {
#pragma clang fp contract(off)
                    mf = (float)(hx + P.hy64 - hall);
                    want = better(mf, pre_key + (uint64_t)c, tau, tau_combo);
                }
            }
            push_candidates(want, mf, pre_key + (uint64_t)c, P, st);
        }
    }
}

template <int K, bool HIST, bool ALL>
__global__ void __launch_bounds__(kBlock) k_scan_km(const DevData D, const ScanParams P, TopState *st)
{
    __shared__ uint32_t lds_hist[HIST ? kHistBins : 1];
    hist_begin<HIST>(lds_hist);
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    km_unit<K, HIST, ALL>(D, P, st, P.j0 + (int64_t)blockIdx.x * (kBlock / kWave) + wid, P.j1, INT32_MAX, 0,
                          lds_hist);
    hist_end<HIST>(lds_hist, P.hist);  // every wave of the block reaches the barrier
}

// ---------------------------------------------------------------------------
// Marginal scans: only the 2^k cells whose genotypes are all 0 or 1 are
// counted by AND + popcount; every other cell follows exactly, by
// inclusion-exclusion, from counts of fewer SNPs (single-row popcounts, the
// pair table, the class total).  Valid when every SNP's three rows are
// disjoint and cover the same samples (checked at upload).
//
// A wavefront takes PW consecutive prefixes of the distribution; the 64
// lanes' last-SNP words are loaded once per quad and shared by all PW
// prefixes, whose rows are wave-uniform scalar loads.
// ---------------------------------------------------------------------------

// In-place transform of one class's 3^3 table from the marginal basis
// (index 2 = "any genotype") to genotype 2, dimension by dimension.
template <class T> __device__ __forceinline__ void mobius3(T *m)
{
#pragma unroll
    for (int ij = 0; ij < 9; ++ij) m[ij * 3 + 2] -= m[ij * 3 + 0] + m[ij * 3 + 1];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) m[i * 9 + 6 + k] -= m[i * 9 + k] + m[i * 9 + 3 + k];
#pragma unroll
    for (int jk = 0; jk < 9; ++jk) m[18 + jk] -= m[jk] + m[9 + jk];
}

template <class T> __device__ __forceinline__ void mobius2(T *m)
{
#pragma unroll
    for (int i = 0; i < 3; ++i) m[i * 3 + 2] -= m[i * 3 + 0] + m[i * 3 + 1];
#pragma unroll
    for (int k = 0; k < 3; ++k) m[6 + k] -= m[k] + m[3 + k];
}

// Order 2.  TABLE = true writes the 8 pair counts (f(x,y), x,y in {0,1}, per
// class) of every pair instead of scoring it.
// One wavefront's unit: prefixes g0 .. g0+PW-1 against tiles t_hi down to
// max(t_lo, first tile past the smallest prefix).
template <int PW, bool TABLE, bool HIST>
__device__ __forceinline__ void scan2m_unit(const DevData &D, const ScanParams &P, TopState *st,
                                            uint32_t *pair_out, int64_t g0, int64_t j1, int t_hi, int t_lo,
                                            uint32_t *lds_hist)
{
    const uint4 *__restrict__ bits = D.bits;
    const int lane = threadIdx.x & 63;
    int pa[PW];
    bool pv[PW];
    int amin = D.n_pad;
#pragma unroll
    for (int i = 0; i < PW; ++i) {
        int tmp[1];
        pv[i] = prefix_at<1>(P, D.n, g0 + i, j1, tmp);
        pa[i] = pv[i] ? tmp[0] : D.n_pad;
        amin = min(amin, pa[i]);
    }
    if (!pv[0]) return;  // slots fill in order: slot 0 empty means all are
    // an empty slot counts a copy of prefix 0 (discarded) so loads stay in range
#pragma unroll
    for (int i = 1; i < PW; ++i)
        if (!pv[i]) pa[i] = pa[0];
    const float tau = TABLE ? 0.0f : st->tau;
    const uint64_t tau_combo = TABLE ? 0 : st->tau_combo;
    const int np = D.n_pad;
    const int t_end = max(t_lo, (amin + 1) >> 6);
    for (int t = min(t_hi, (np >> 6) - 1); t >= t_end; --t) {  // last tile first (shared in L2)
        const int c = (t << 6) | lane;
        const int tile_last = (t << 6) | 63;
        uint32_t f[PW][2][4];
        bool live[PW];
#pragma unroll
        for (int i = 0; i < PW; ++i) {
            live[i] = pv[i] && tile_last > pa[i];
#pragma unroll
            for (int x = 0; x < 8; ++x) f[i][x >> 2][x & 3] = 0;
        }
#pragma unroll
        for (int cls = 0; cls < 2; ++cls) {
            const int nq = D.quads[cls];
            if (nq <= 0) continue;
            const uint4 *row = bits + D.class_off[cls];
            const int voff = c * 16;
            // same issue order as the order-3 scan: prefix i+1's rows (two
            // register sets, alternately) before prefix i is counted, the
            // lanes' next quad before the last two prefixes of this one
            __amdgpu_buffer_rsrc_t r = quad_rsrc(row, np);
            uint4 c0 = ld_lane(r, np, 0, voff), c1 = ld_lane(r, np, 1, voff);
            uint4 X[2], Y[2];
            X[0] = ld_uniform(r, np, 0, pa[0]);
            X[1] = ld_uniform(r, np, 1, pa[0]);
            for (int q = 0; q < nq; ++q) {
                const uint4 *nrow = q + 1 < nq ? row + (size_t)3 * np : row;
                const __amdgpu_buffer_rsrc_t nr = quad_rsrc(nrow, np);
                uint4 n0, n1;
#pragma unroll
                for (int i = 0; i < PW; ++i) {
                    if (i == PW - 2 || (PW == 1 && i == 0)) {
                        n0 = ld_lane(nr, np, 0, voff);
                        n1 = ld_lane(nr, np, 1, voff);
                    }
                    const __amdgpu_buffer_rsrc_t r2 = i + 1 < PW ? r : nr;
                    const int ia = pa[(i + 1) % PW];
                    uint4 *nxt = (i & 1) ? X : Y;
                    nxt[0] = ld_uniform(r2, np, 0, ia);
                    nxt[1] = ld_uniform(r2, np, 1, ia);
                    __builtin_amdgcn_sched_barrier(0);
                    const uint4 *cur = (i & 1) ? Y : X;
                    if (live[i]) {  // uniform
#pragma unroll
                        for (int w = 0; w < 4; ++w) {
                            const uint32_t x0 = comp(c0, w), x1 = comp(c1, w);
                            const uint32_t p0 = comp(cur[0], w), p1 = comp(cur[1], w);
                            and2_bcnt4(f[i][cls][0], f[i][cls][1], f[i][cls][2], f[i][cls][3], p0, x0, p0, x1, p1,
                                       x0, p1, x1);
                        }
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
                c0 = n0;
                c1 = n1;
                row = nrow;
                r = nr;
            }
        }
        const bool in_range = c < D.n;
        uint32_t sc[2][2] = {{0, 0}, {0, 0}};
        if (!TABLE && in_range) {
            const uint32_t *sp = D.singles + (size_t)c * 6;
            sc[0][0] = sp[0], sc[0][1] = sp[1], sc[1][0] = sp[3], sc[1][1] = sp[4];
        }
#pragma unroll
        for (int i = 0; i < PW; ++i) {
            if (!live[i]) continue;  // uniform
            const bool valid = in_range && c > pa[i];
            if (TABLE) {
                if (valid) {
                    uint4 *o = (uint4 *)(pair_out + pair_index(pa[i], c, D.n) * 8);
                    o[0] = make_uint4(f[i][0][0], f[i][0][1], f[i][0][2], f[i][0][3]);
                    o[1] = make_uint4(f[i][1][0], f[i][1][1], f[i][1][2], f[i][1][3]);
                }
                continue;
            }
            uint32_t m0[9], m1[9];
            const uint32_t *sa = D.singles + (size_t)pa[i] * 6;
#pragma unroll
            for (int cls = 0; cls < 2; ++cls) {
                uint32_t *m = cls ? m1 : m0;
                m[0] = f[i][cls][0];
                m[1] = f[i][cls][1];
                m[3] = f[i][cls][2];
                m[4] = f[i][cls][3];
                m[2] = sa[cls * 3 + 0];
                m[5] = sa[cls * 3 + 1];
                m[6] = sc[cls][0];
                m[7] = sc[cls][1];
                m[8] = D.total[cls];
                mobius2(m);
            }
            score_or_sample<HIST, 9>(
                valid, m0, m1, [&] { return prefix_key(P.binom, D.n, &pa[i], 1) + (uint64_t)c; }, tau, tau_combo, P,
                st, lds_hist);
        }
    }
}

template <int PW, bool TABLE, bool HIST>
__global__ void __launch_bounds__(kBlock)
k_scan2m(const DevData D, const ScanParams P, TopState *st, uint32_t *pair_out)
{
    __shared__ uint32_t lds_hist[HIST ? kHistBins : 1];
    hist_begin<HIST>(lds_hist);
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // gridDim.y > 1: the tiles of last SNPs split in gridDim.y chunks (more
    // wavefronts for the pair table of a small dataset)
    int t_hi = INT32_MAX, t_lo = 0;
    if (gridDim.y > 1) {
        const int tiles = ((D.n_pad >> 6) + (int)gridDim.y - 1) / (int)gridDim.y;
        t_hi = (D.n_pad >> 6) - 1 - (int)blockIdx.y * tiles;
        t_lo = t_hi - tiles + 1;
    }
    scan2m_unit<PW, TABLE, HIST>(D, P, st, pair_out, P.j0 + ((int64_t)blockIdx.x * (kBlock / kWave) + wid) * PW,
                                 P.j1, t_hi, t_lo, lds_hist);
    hist_end<HIST>(lds_hist, P.hist);
}

// Order 3.  A wavefront counts PW prefixes (a, b) against the 64 SNPs of a
// tile.  Every cell word is one v_bitop3_b32 (a_x & b_y & c_z) and one
// accumulating v_bcnt_u32_b32.  All rows arrive by buffer loads, which
// complete in order: the rows of prefix i+1 are issued before prefix i is
// counted, and the lanes' next quad before the last two prefixes, so every
// wait covers data that had a whole prefix of counting to arrive.
// Unit: prefixes g0 .. g0+PW-1, tiles t_hi down to max(t_lo, first tile past
// the smallest prefix).
template <int PW, bool HIST>
__device__ __forceinline__ void scan3m_unit(const DevData &D, const ScanParams &P, TopState *st, int64_t g0,
                                            int64_t j1, int t_hi, int t_lo, uint32_t *lds_hist)
{
    const uint4 *__restrict__ bits = D.bits;
    const int lane = threadIdx.x & 63;
    int pa[PW], pb[PW];
    bool pv[PW];
    int bmin = D.n_pad;
#pragma unroll
    for (int i = 0; i < PW; ++i) {
        int tmp[2];
        pv[i] = prefix_at<2>(P, D.n, g0 + i, j1, tmp);
        pa[i] = pv[i] ? tmp[0] : 0;
        pb[i] = pv[i] ? tmp[1] : D.n_pad;
        bmin = min(bmin, pb[i]);
    }
    if (!pv[0]) return;  // slots fill in order: slot 0 empty means all are
    // an empty slot counts a copy of prefix 0 (discarded) so loads stay in range
#pragma unroll
    for (int i = 1; i < PW; ++i)
        if (!pv[i]) pa[i] = pa[0], pb[i] = pb[0];
    const float tau = st->tau;
    const uint64_t tau_combo = st->tau_combo;
    const int np = D.n_pad, n = D.n;
    const uint32_t *pair = P.pair;
    const int t_end = max(t_lo, (bmin + 1) >> 6);
    for (int t = min(t_hi, (np >> 6) - 1); t >= t_end; --t) {  // last tile first (shared in L2)
        const int c = (t << 6) | lane;
        const int tile_last = (t << 6) | 63;
        uint32_t f[PW][2][8];
        bool live[PW];
#pragma unroll
        for (int i = 0; i < PW; ++i) {
            live[i] = pv[i] && tile_last > pb[i];
#pragma unroll
            for (int x = 0; x < 16; ++x) f[i][x >> 3][x & 7] = 0;
        }
#pragma unroll
        for (int cls = 0; cls < 2; ++cls) {
            const int nq = D.quads[cls];
            if (nq <= 0) continue;
            const uint4 *row = bits + D.class_off[cls];
            const int voff = c * 16;
            __amdgpu_buffer_rsrc_t r = quad_rsrc(row, np);
            // Loads complete in order, so the issue order decides what each
            // wait covers: the rows of prefix i+1 (two register sets used
            // alternately) are issued before prefix i is counted, and the
            // lanes' next quad before the last two prefixes of this one.
            uint4 c0 = ld_lane(r, np, 0, voff), c1 = ld_lane(r, np, 1, voff);
            uint4 X[4], Y[4];
            X[0] = ld_uniform(r, np, 0, pa[0]);
            X[1] = ld_uniform(r, np, 1, pa[0]);
            X[2] = ld_uniform(r, np, 0, pb[0]);
            X[3] = ld_uniform(r, np, 1, pb[0]);
            for (int q = 0; q < nq; ++q) {
                const uint4 *nrow = q + 1 < nq ? row + (size_t)3 * np : row;
                const __amdgpu_buffer_rsrc_t nr = quad_rsrc(nrow, np);
                uint4 n0, n1;
#pragma unroll
                for (int i = 0; i < PW; ++i) {
                    if (i == PW - 2 || (PW == 1 && i == 0)) {
                        n0 = ld_lane(nr, np, 0, voff);
                        n1 = ld_lane(nr, np, 1, voff);
                    }
                    const __amdgpu_buffer_rsrc_t r2 = i + 1 < PW ? r : nr;
                    const int ia = pa[(i + 1) % PW], ib = pb[(i + 1) % PW];
                    uint4 *nxt = (i & 1) ? X : Y;
                    nxt[0] = ld_uniform(r2, np, 0, ia);
                    nxt[1] = ld_uniform(r2, np, 1, ia);
                    nxt[2] = ld_uniform(r2, np, 0, ib);
                    nxt[3] = ld_uniform(r2, np, 1, ib);
                    __builtin_amdgcn_sched_barrier(0);
                    const uint4 *cur = (i & 1) ? Y : X;
                    if (live[i]) {  // uniform: the tile holds candidates for prefix i
#pragma unroll
                        for (int w = 0; w < 4; ++w) {
                            const uint32_t x0 = comp(c0, w), x1 = comp(c1, w);
                            const uint32_t u0 = comp(cur[0], w), u1 = comp(cur[1], w);
                            const uint32_t v0 = comp(cur[2], w), v1 = comp(cur[3], w);
                            and3_bcnt4(f[i][cls][0], f[i][cls][1], f[i][cls][2], f[i][cls][3], u0, v0, x0, u0, v0,
                                       x1, u0, v1, x0, u0, v1, x1);
                            and3_bcnt4(f[i][cls][4], f[i][cls][5], f[i][cls][6], f[i][cls][7], u1, v0, x0, u1, v0,
                                       x1, u1, v1, x0, u1, v1, x1);
                        }
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
                c0 = n0;
                c1 = n1;
                row = nrow;
                r = nr;
            }
        }
        const bool in_range = c < n;
        uint32_t sc[2][2] = {{0, 0}, {0, 0}};
        if (in_range) {
            const uint32_t *sp = D.singles + (size_t)c * 6;
            sc[0][0] = sp[0], sc[0][1] = sp[1], sc[1][0] = sp[3], sc[1][1] = sp[4];
        }
#pragma unroll
        for (int i = 0; i < PW; ++i) {
            if (!pv[i] || tile_last <= pb[i]) continue;  // uniform
            const bool valid = in_range && c > pb[i];
            uint32_t fac[8] = {0, 0, 0, 0, 0, 0, 0, 0}, fbc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            if (valid) {
                const uint4 *pac = (const uint4 *)(pair + pair_index(pa[i], c, n) * 8);
                const uint4 *pbc = (const uint4 *)(pair + pair_index(pb[i], c, n) * 8);
                const uint4 u0 = pac[0], u1 = pac[1], v0 = pbc[0], v1 = pbc[1];
                fac[0] = u0.x, fac[1] = u0.y, fac[2] = u0.z, fac[3] = u0.w;
                fac[4] = u1.x, fac[5] = u1.y, fac[6] = u1.z, fac[7] = u1.w;
                fbc[0] = v0.x, fbc[1] = v0.y, fbc[2] = v0.z, fbc[3] = v0.w;
                fbc[4] = v1.x, fbc[5] = v1.y, fbc[6] = v1.z, fbc[7] = v1.w;
            }
            const uint32_t *pab = pair + pair_index(pa[i], pb[i], n) * 8;
            const uint32_t *sa = D.singles + (size_t)pa[i] * 6;
            const uint32_t *sb = D.singles + (size_t)pb[i] * 6;
            uint32_t m0[27], m1[27];
#pragma unroll
            for (int cls = 0; cls < 2; ++cls) {
                uint32_t *m = cls ? m1 : m0;
                // m[x*9 + y*3 + z], x,y,z in {0, 1, 2 = any}
#pragma unroll
                for (int x = 0; x < 2; ++x)
#pragma unroll
                    for (int y = 0; y < 2; ++y) {
                        m[x * 9 + y * 3 + 0] = f[i][cls][(x * 2 + y) * 2 + 0];
                        m[x * 9 + y * 3 + 1] = f[i][cls][(x * 2 + y) * 2 + 1];
                        m[x * 9 + y * 3 + 2] = pab[cls * 4 + x * 2 + y];
                    }
#pragma unroll
                for (int x = 0; x < 2; ++x)
#pragma unroll
                    for (int z = 0; z < 2; ++z) m[x * 9 + 6 + z] = fac[cls * 4 + x * 2 + z];
#pragma unroll
                for (int y = 0; y < 2; ++y)
#pragma unroll
                    for (int z = 0; z < 2; ++z) m[18 + y * 3 + z] = fbc[cls * 4 + y * 2 + z];
                m[0 * 9 + 8] = sa[cls * 3 + 0];
                m[1 * 9 + 8] = sa[cls * 3 + 1];
                m[18 + 0 * 3 + 2] = sb[cls * 3 + 0];
                m[18 + 1 * 3 + 2] = sb[cls * 3 + 1];
                m[18 + 6 + 0] = sc[cls][0];
                m[18 + 6 + 1] = sc[cls][1];
                m[26] = D.total[cls];
                mobius3(m);
            }
            auto key = [&] {
                const int pre[2] = {pa[i], pb[i]};
                return prefix_key(P.binom, n, pre, 2) + (uint64_t)c;
            };
            score_or_sample<HIST, 27>(valid, m0, m1, key, tau, tau_combo, P, st, lds_hist);
        }
    }
}

template <int PW, bool HIST>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3)))   // <= 168 registers
k_scan3m(const DevData D, const ScanParams P, TopState *st)
{
    __shared__ uint32_t lds_hist[HIST ? kHistBins : 1];
    hist_begin<HIST>(lds_hist);
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    scan3m_unit<PW, HIST>(D, P, st, P.j0 + ((int64_t)blockIdx.x * (kBlock / kWave) + wid) * PW, P.j1, INT32_MAX,
                          0, lds_hist);
    hist_end<HIST>(lds_hist, P.hist);
}

// ---------------------------------------------------------------------------
// Order 2 on the matrix cores.  With one {0,1} indicator row per SNP and
// genotype (I_g[s][k] = 1 iff sample k of the class has genotype g at SNP s),
// the four marginal pair counts are dot products over the samples:
// f(x, y)[a][b] = sum_k I_x[a][k] I_y[b][k] -- a symmetric GEMM per class; the
// other five cells follow from the singles and the class total exactly as in
// the popcount scan.  Two operand formats, both exact:
//   int8: v_mfma_i32_32x32x32_i8, int32 accumulation;
//   fp4 (e2m1, 1.0 = nibble 0x2): v_mfma_f32_32x32x64_f8f6f4, fp32
//     accumulation of 0/1 products -- exact while a count stays below 2^24
//     (checked on the host); half the bytes per product and twice the rate.
// ---------------------------------------------------------------------------
// f(integral_constant<int, 0>) ... f(integral_constant<int, N - 1>), inline.
template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F> __device__ __forceinline__ void static_for(F &&f)
{
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef float v16f __attribute__((ext_vector_type(16)));

constexpr int kMfmaTile = 64;       // SNPs per tile side (per workgroup); a stage is one
                                    // 128-sample quad per row
constexpr int kMfmaSuper = 8;       // tiles per side of the super-blocks run per XCD

// LDS row pitch per format: the data (128 int8 or 64 bytes of fp4) + 16, so
// that the 16 rows one ds_read_b128 phase touches hit distinct banks.
template <bool FP4> struct MfmaFmt {
    static constexpr int row = FP4 ? 80 : 144;
    static constexpr int substeps = FP4 ? 2 : 4;   // MFMAs per 128-sample stage and accumulator
};

// Direct-to-LDS buffer load of 16 bytes per lane from byte offset voff of
// the resource: lane l's bytes land at lds_group + 16 l (lds_group
// wave-uniform).  The per-lane offsets stay fixed while the resource base
// moves (scalar arithmetic only).
__device__ __forceinline__ void dma16(__amdgpu_buffer_rsrc_t r, uint32_t voff, int8_t *lds_group)
{
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (__attribute__((address_space(3))) void *)lds_group, 16, voff, 0, 0,
                                            0);
}

// Wait until at most N vector-memory operations of this wave are pending
// (vmcnt N; lgkm and exp counters not waited on).
template <int N> __device__ __forceinline__ void wait_vmcnt()
{
    static_assert(N >= 0 && N < 16, "vmcnt immediate");
    __builtin_amdgcn_s_waitcnt(0x0f70 | N);
}

// wait_vmcnt<L * s>() for a runtime, wave-uniform s (clamped to 0 .. S).
template <int S, int L> __device__ __forceinline__ void wait_stages(int s)
{
    if constexpr (S > 0) {
        if (s >= S) {
            wait_vmcnt<L * S>();
            return;
        }
        wait_stages<S - 1, L>(s);
    } else {
        wait_vmcnt<0>();
    }
}

// Workgroup barrier with no memory fence (a fence would drain every pending
// direct-to-LDS load, including the ones meant to stay in flight); the
// waits are explicit, the empty asm statements keep the compiler from moving
// memory accesses across the barrier.
__device__ __forceinline__ void stage_barrier()
{
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// Eight bits (b < 256) as eight e2m1 nibbles, bit i -> nibble i = 0x2 (1.0)
// or 0x0: the bits are spread 4 apart by three shift-or-mask steps, then
// moved to bit 1 of each nibble.
__device__ __forceinline__ uint32_t byte_fp4(uint32_t b)
{
    b = (b | (b << 12)) & 0x000F000Fu;
    b = (b | (b << 6)) & 0x03030303u;
    b = (b | (b << 3)) & 0x11111111u;
    return b << 1;
}

// fp4 image of a dataset's genotype rows 0 and 1 (built once per dataset):
// the 128 samples of quad q, genotype g, SNP s of a class are 64 bytes = 4
// uint4 at class_off[c] + 4 ((2 q + g) n_pad + s), nibble j = sample j, so
// the matrix-core scans copy operand rows instead of expanding bits.
__global__ void k_fp4_image(const uint4 *__restrict__ bits, DevData D, uint4 *__restrict__ out, uint64_t off0,
                            uint64_t off1)
{
    const uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t per0 = 2ull * D.quads[0] * D.n_pad, total = per0 + 2ull * D.quads[1] * D.n_pad;
    if (id >= total) return;
    const int cls = id < per0 ? 0 : 1;
    const uint64_t v = cls ? id - per0 : id;
    const int s = (int)(v % D.n_pad);
    const uint64_t qg = v / D.n_pad;
    const int g = (int)(qg & 1), q = (int)(qg >> 1);
    const uint4 w = bits[D.class_off[cls] + ((uint64_t)q * 3 + g) * D.n_pad + s];
    uint4 *dst = out + (cls ? off1 : off0) + 4 * v;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const uint32_t x = comp(w, e);
        dst[e] = make_uint4(byte_fp4(x & 255u), byte_fp4(x >> 8 & 255u), byte_fp4(x >> 16 & 255u), byte_fp4(x >> 24));
    }
}


// Four bits of v (from bit `lo`) as four bytes 0/1: n * 0x00204081 puts bit i
// of the nibble n at bit 8i (the shifted copies do not overlap), the mask
// keeps those bits.
__device__ __forceinline__ uint32_t nibble_bytes(uint32_t v, int lo)
{
    return ((v >> lo & 15u) * 0x00204081u) & 0x01010101u;
}

// 128 bits of a quad as 128 indicator bytes in one LDS row: bit j of 32-bit
// word e becomes byte 32 e + j.  Rows of both operands use the same order, so
// dot products over the samples are unchanged.
__device__ __forceinline__ void expand_quad(const uint4 &q, int8_t *row)
{
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const uint32_t v = comp(q, e);
#pragma unroll
        for (int h = 0; h < 2; ++h)
            *(uint4 *)(row + 32 * e + 16 * h) = make_uint4(nibble_bytes(v, 16 * h), nibble_bytes(v, 16 * h + 4),
                                                         nibble_bytes(v, 16 * h + 8), nibble_bytes(v, 16 * h + 12));
    }
}

// Tile u of the upper triangle of the T x T grid of 64-SNP tiles, row-major
// (row ta holds tiles ta..T-1).
__device__ __forceinline__ void tri_tile(int64_t u, int T, int &ta, int &tb)
{
    // ta = largest r with r*T - r(r-1)/2 <= u
    const double tt = (double)T + 0.5;
    int r = (int)(tt - sqrt(tt * tt - 2.0 * (double)u));
    r = max(0, min(r, T - 1));
    while (r > 0 && (int64_t)r * T - (int64_t)r * (r - 1) / 2 > u) --r;
    while (r + 1 < T && (int64_t)(r + 1) * T - (int64_t)(r + 1) * r / 2 <= u) ++r;
    ta = r;
    tb = (int)(u - ((int64_t)r * T - (int64_t)r * (r - 1) / 2)) + r;
}

// One workgroup = one 64 x 64 tile (4 waves, 32 x 32 SNPs each).  The
// launch covers super-blocks [s0, s1) of the triangle of S x S super-blocks
// (8 x 8 tiles each): block v runs on XCD v % 8, and the 64 blocks an XCD
// holds at once are one super-block, whose 8 + 8 row panels they share in L2.
// Each stage covers one 128-sample quad of 256 operand rows (A: 64 SNPs x
// genotypes 0/1, B likewise).  fp4: the rows are copied from the dataset's
// fp4 image (four 16-byte units per thread); int8: each thread loads one
// 128-bit quad of the bit tables and expands it to bytes in LDS.
// LDS of one k_mfma2 workgroup, [buffer][A/B][row][k]: fp4, a ring of
// kM2Ring stages of 64-byte rows filled by direct-to-LDS loads; int8, two
// buffers of expanded rows.
constexpr int kM2Ring = 4;
template <bool FP4> struct Mfma2LdsT {
    typedef int8_t type[2][2][2 * kMfmaTile][MfmaFmt<false>::row];
};
template <> struct Mfma2LdsT<true> {
    typedef int8_t type[kM2Ring][2][2 * kMfmaTile][64];
};
template <bool FP4> using Mfma2Lds = typename Mfma2LdsT<FP4>::type;
// Segment s (16 bytes) of LDS row r (fp4 ring) is stored at s ^ m2_swz(r):
// the 32-row fragment reads (lane l: row l % 32, segment 2 k + l / 32) are
// then free of bank conflicts (measured with a ds_read_b128 probe: 4.7
// cycles per wave per CU, as a linear read).
__device__ __forceinline__ int m2_swz(int r) { return (r >> 2) & 3; }

// One workgroup's work item: tile `slot` (of 64) of super-block su; returns at
// once for tiles outside the triangle.  Every wave of the workgroup calls it
// with the same arguments.
template <bool FP4>
__device__ __forceinline__ void mfma2_tile(const DevData &D, const ScanParams &P, TopState *st, const Fp4Image &F4,
                                           int64_t su, int slot, int T, Mfma2Lds<FP4> &lds)
{
    typedef MfmaFmt<FP4> F;
    typedef typename std::conditional<FP4, v16f, v16i>::type Acc;
    const int S = (T + kMfmaSuper - 1) / kMfmaSuper;
    int sa, sc;
    tri_tile(su, S, sa, sc);
    const int ta = sa * kMfmaSuper + (slot >> 3), tb = sc * kMfmaSuper + (slot & 7);
    if (ta > tb || tb >= T) return;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6), wa = w >> 1, wb = w & 1;
    const int np = D.n_pad;
    const int a0 = ta * kMfmaTile, b0 = tb * kMfmaTile;
    Acc acc[2][2][2];
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int g = 0; g < 2; ++g)
#pragma unroll
            for (int h = 0; h < 2; ++h) acc[c][g][h] = (Acc){0};
    if constexpr (FP4) {
        // staging: each wave issues 4 direct-to-LDS loads per stage, rows
        // 16 w .. 16 w + 15 of the A genotype-0, A genotype-1, B genotype-0
        // and B genotype-1 panels; lane l fills row 16 w + l / 4, stored
        // segment l % 4, so it loads the row's segment (l % 4) ^ m2_swz(row).
        // Byte offsets within a quad of the image (< 128 n_pad).
        const int lr = lane >> 2, sg = (lane & 3) ^ m2_swz(lr);
        const uint32_t oa0 = 16u * (uint32_t)(4 * (a0 + 16 * w + lr) + sg);
        const uint32_t oa1 = 16u * (uint32_t)(4 * (np + a0 + 16 * w + lr) + sg);
        const uint32_t ob0 = 16u * (uint32_t)(4 * (b0 + 16 * w + lr) + sg);
        const uint32_t ob1 = 16u * (uint32_t)(4 * (np + b0 + 16 * w + lr) + sg);
        const int r = lane & 31;
#pragma unroll
        for (int cls = 0; cls < 2; ++cls) {
            const int iters = D.quads[cls];
            const uint4 *img = F4.base + F4.class_off[cls];   // quad q at + q * 8 * np
            auto issue = [&](int q) {
                const __amdgpu_buffer_rsrc_t rq =
                    __builtin_amdgcn_make_buffer_rsrc((void *)(img + (int64_t)q * 8 * np), (short)0, 128 * np, 0x00020000);
                auto &sl = lds[q % kM2Ring];
                dma16(rq, oa0, &sl[0][16 * w][0]);
                dma16(rq, oa1, &sl[0][kMfmaTile + 16 * w][0]);
                dma16(rq, ob0, &sl[1][16 * w][0]);
                dma16(rq, ob1, &sl[1][kMfmaTile + 16 * w][0]);
            };
#pragma unroll
            for (int q = 0; q < kM2Ring - 1; ++q)
                if (q < iters) issue(q);
            // the stages with at least kM2Ring - 1 stages after them run a
            // body without branches (constant wait, always one stage to
            // issue), so that the fragment reads and the products share a
            // basic block and each product waits only for its own operands;
            // the last kM2Ring - 1 stages run the general body
            auto stage = [&](int it, bool steady) __attribute__((always_inline)) {
                // this wave's loads of stage it are complete once at most those
                // of the later stages issued (4 each) are pending; the barrier
                // publishes every wave's part
                if (steady) wait_vmcnt<4 * (kM2Ring - 2)>();
                else wait_stages<kM2Ring - 2, 4>(iters - 1 - it);
                stage_barrier();
                const auto &sl = lds[it % kM2Ring];
                v4i fa[2][2], fb[2][2];   // [substep][genotype]: 32 fp4 samples of row r per lane
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    const int ko = ((2 * s + (lane >> 5)) ^ m2_swz(r)) << 4;
#pragma unroll
                    for (int g = 0; g < 2; ++g) {
                        fa[s][g] = *(const v4i *)&sl[0][g * kMfmaTile + wa * 32 + r][ko];
                        fb[s][g] = *(const v4i *)&sl[1][g * kMfmaTile + wb * 32 + r][ko];
                    }
                }
                // every wave is past stage it - 1, whose slot takes stage it + kM2Ring - 1
                if (steady || it + kM2Ring - 1 < iters) issue(it + kM2Ring - 1);
#pragma unroll
                for (int s = 0; s < 2; ++s)
#pragma unroll
                    for (int g = 0; g < 2; ++g)
#pragma unroll
                        for (int h = 0; h < 2; ++h) {
                            const v8i a8 = {fa[s][g][0], fa[s][g][1], fa[s][g][2], fa[s][g][3], 0, 0, 0, 0};
                            const v8i b8 = {fb[s][h][0], fb[s][h][1], fb[s][h][2], fb[s][h][3], 0, 0, 0, 0};
                            // formats 4/4 = e2m1; scale operands 0 select the unscaled product
                            acc[cls][g][h] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a8, b8, acc[cls][g][h],
                                                                                           4, 4, 0, 0, 0, 0);
                        }
            };
            int it = 0;
            for (; it + kM2Ring - 1 < iters; ++it) stage(it, true);
            for (; it < iters; ++it) stage(it, false);
            // the next class's first loads reuse the ring
            stage_barrier();
        }
    } else {
        // int8 staging: this thread's row, operand m (A / B), genotype g, SNP;
        // each stage loads one 128-bit quad per row and expands it to bytes
        const int m_ld = tid >> 7, row_ld = tid & 127, g_ld = row_ld >> 6;
        const int snp_ld = (m_ld ? b0 : a0) + (row_ld & 63);
#pragma unroll
        for (int cls = 0; cls < 2; ++cls) {
            const int iters = D.quads[cls];
            const uint4 *src = D.bits + D.class_off[cls] + (uint64_t)g_ld * np + snp_ld;   // quad q at + q * 3 * np
            // quads it+1 and it+2 wait in registers, quad it+3 is loaded
            // during stage it: two stages of latency hiding for every load
            uint4 q1 = src[0], q2 = make_uint4(0, 0, 0, 0);
            expand_quad(q1, lds[0][m_ld][row_ld]);
            if (iters > 1) q1 = src[(uint64_t)3 * np];
            if (iters > 2) q2 = src[(uint64_t)6 * np];
            __syncthreads();
            for (int it = 0; it < iters; ++it) {
                const int buf = it & 1;
                uint4 q3 = make_uint4(0, 0, 0, 0);
                if (it + 3 < iters) q3 = src[(uint64_t)(it + 3) * 3 * np];
#pragma unroll
                for (int s = 0; s < F::substeps; ++s) {
                    // 16 operand bytes per lane: 16 int8 samples of row r
                    const int kk = s * 32 + (lane >> 5) * 16, r = lane & 31;
                    v4i fa[2], fb[2];
#pragma unroll
                    for (int g = 0; g < 2; ++g) {
                        fa[g] = *(const v4i *)&lds[buf][0][g * kMfmaTile + wa * 32 + r][kk];
                        fb[g] = *(const v4i *)&lds[buf][1][g * kMfmaTile + wb * 32 + r][kk];
                    }
#pragma unroll
                    for (int g = 0; g < 2; ++g)
#pragma unroll
                        for (int h = 0; h < 2; ++h)
                            acc[cls][g][h] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[g], fb[h], acc[cls][g][h], 0, 0, 0);
                }
                if (it + 1 < iters) expand_quad(q1, lds[buf ^ 1][m_ld][row_ld]);
                __syncthreads();
                q1 = q2;
                q2 = q3;
            }
        }
    }
    // epilogue: lane holds column b = b0 + wb*32 + lane%32 and rows
    // a = a0 + wa*32 + 8r + 4(lane/32) + q, r, q < 4
    const float tau = st->tau;
    const uint64_t tau_combo = st->tau_combo;
    const int b = b0 + wb * 32 + (lane & 31);
    const bool b_ok = b < D.n;
    uint32_t sb[2][2] = {{0, 0}, {0, 0}};
    if (b_ok) {
        const uint32_t *sp = D.singles + (size_t)b * 6;
        sb[0][0] = sp[0], sb[0][1] = sp[1], sb[1][0] = sp[3], sb[1][1] = sp[4];
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = r * 4 + q;
            const int a = a0 + wa * 32 + 8 * r + 4 * (lane >> 5) + q;
            const int64_t rel = (int64_t)a - P.offset;
            const bool valid = b_ok && a < b && rel >= 0 && (P.step == 1 || rel % P.step == 0);
            const uint32_t *sa = D.singles + (size_t)a * 6;
            // counts as FP32 (fp4 products) or uint32_t (int8 products): both
            // exact, and the inclusion-exclusion stays exact in either
            typedef typename std::conditional<FP4, float, uint32_t>::type Cnt;
            Cnt m0[9], m1[9];
#pragma unroll
            for (int cls = 0; cls < 2; ++cls) {
                Cnt *m = cls ? m1 : m0;
                m[0] = (Cnt)acc[cls][0][0][i];
                m[1] = (Cnt)acc[cls][0][1][i];
                m[3] = (Cnt)acc[cls][1][0][i];
                m[4] = (Cnt)acc[cls][1][1][i];
                m[2] = (Cnt)sa[cls * 3 + 0];
                m[5] = (Cnt)sa[cls * 3 + 1];
                m[6] = (Cnt)sb[cls][0];
                m[7] = (Cnt)sb[cls][1];
                m[8] = (Cnt)D.total[cls];
                mobius2(m);
            }
            score_and_push<9>(valid, m0, m1, [&] { return prefix_key(P.binom, D.n, &a, 1) + (uint64_t)b; }, tau,
                              tau_combo, P, st);
        }
}

template <bool FP4>
__global__ void __launch_bounds__(kBlock)
k_mfma2(const DevData D, const ScanParams P, TopState *st, const Fp4Image F4, int64_t s0, int64_t s1, int T)
{
    __shared__ __attribute__((aligned(16))) Mfma2Lds<FP4> lds;
    const int64_t k = blockIdx.x >> 3, su = s0 + (k >> 6) * 8 + (blockIdx.x & 7);
    if (su >= s1) return;
    mfma2_tile<FP4>(D, P, st, F4, su, (int)(k & 63), T, lds);
}

// Wide order-2 tiles: one workgroup of AT * 4 waves takes AT * 64 a x 128 b
// (the a tiles AT j .. and the b tiles 2 i, 2 i + 1 of a super-block), AT = 1
// (k_mfma2w, 64 x 128) or 2 (k_mfma2x, 128 x 128); wave w multiplies the 64 a
// rows a0 + 64 (w / 4) ... by the 32 b rows b0 + 32 (w % 4) ...: 16 products
// per 128-sample stage against 12 fragment reads (8 per 8 in mfma2_tile), and
// 24 KB (AT = 1) or 32 KB (AT = 2) staged per 16 AT * 4 products (16 KB per
// 32).  One class's counts fill the 128 accumulators; both classes wait as
// 16-bit pairs in private memory (every class holds fewer than 2^16 samples),
// the epilogue scores them in two passes as mfma3w_tile does.  Stage layout:
// (2 AT + 4) * 64 rows of 64 bytes, rows g * 64 AT + i the a rows (genotype
// g, SNP a0 + i), rows 128 AT + h * 128 + j the b rows; segments swizzled as
// in mfma2_tile.
constexpr int kM2wRing = 3;
// the 128 x 128 tile runs one workgroup per CU: a deeper ring (128 KB)
constexpr int kM2xRing = 4;
template <int AT> constexpr int m2w_stage() { return (128 * AT + 256) * 64; }
constexpr int kM2wStage = m2w_stage<1>();
// Words of the first class's 16 packed count words that stay in registers
// through the second class's stages (the rest wait in private memory).
constexpr int kM2wKeep = 8;
// The same in the batched steps' persistent kernel, whose queue state takes
// registers of its own.
constexpr int kBatchM2wKeep = 4;

// Counts of instance row block (ab, r) as 16-bit pairs: word (g * 2 + h) * 2
// + q / 2 holds f(g, h) of row 8 r + 4 (lane / 32) + q of a block ab, in the
// low (q even) or high half.
__device__ __forceinline__ v4i pack_counts2(const v16f (&acc)[2][2][2], int ab, int r, int half)
{
    auto pk = [](float lo, float hi) __attribute__((always_inline)) {
        return (int)((uint32_t)lo | (uint32_t)hi << 16);
    };
    const int g = half;   // words 4 half .. 4 half + 3: g = half, h = 0, 1
    const v16f &v0 = acc[ab][g][0], &v1 = acc[ab][g][1];
    return (v4i){pk(v0[r * 4 + 0], v0[r * 4 + 1]), pk(v0[r * 4 + 2], v0[r * 4 + 3]), pk(v1[r * 4 + 0], v1[r * 4 + 1]),
                 pk(v1[r * 4 + 2], v1[r * 4 + 3])};
}

template <int KEEP, int AT = 1, int R = kM2wRing>
__device__ __forceinline__ void mfma2w_tile(const DevData &D, const ScanParams &P, TopState *st, const Fp4Image &F4,
                                            int64_t su, int slot, int T, int8_t (*lds)[m2w_stage<AT>()])
{
    static_assert(AT == 1 || AT == 2, "a tiles per workgroup");
    // groups of 16 rows per stage and per wave
    constexpr int kGroups = 8 * AT + 16, kPerWave = kGroups / (4 * AT);
    const int S = (T + kMfmaSuper - 1) / kMfmaSuper;
    int sa, sc;
    tri_tile(su, S, sa, sc);
    const int ta = sa * kMfmaSuper + AT * (slot >> 2), tb = sc * kMfmaSuper + 2 * (slot & 3);
    if (ta > tb + 1 || ta >= T || tb >= T) return;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wa = w >> 2, wb = w & 3;   // the wave's a half (AT = 2) and b quarter
    const int np = D.n_pad;
    const int a0 = ta * kMfmaTile, b0 = tb * kMfmaTile;
    // staging: wave w loads the 16-row groups kPerWave w .. of the kGroups
    // (group G: rows 16 G ..); lane l fills row l / 4 of the group, stored
    // segment l % 4, so it loads the row's segment (l % 4) ^ m2_swz(row).
    // Offsets within a quad of the image; rows past the padded SNPs (an odd
    // last tile) read other rows of the quad or, past its end, zeros (the
    // buffer's range check): the epilogue drops their instances (a < b < n)
    const int lr = lane >> 2, sg = (lane & 3) ^ m2_swz(lr);
    uint32_t vo[kPerWave];
#pragma unroll
    for (int t = 0; t < kPerWave; ++t) {
        const int G = kPerWave * w + t;
        const int gen = G < 8 * AT ? G / (4 * AT) : (G - 8 * AT) >> 3;
        const int snp = G < 8 * AT ? a0 + 16 * (G % (4 * AT)) : b0 + 16 * ((G - 8 * AT) & 7);
        vo[t] = 16u * (uint32_t)(4 * (gen * np + snp + lr) + sg);
    }
    v16f acc[2][2][2];   // [a block][g][h], the class being scanned
    // the first class's counts (pack_counts2 words, block (ab, r), half ->
    // 2 (4 ab + r) + half) in private memory, out of the registers the
    // second class needs; `marked`: both classes' counts (f(g, h) of class 0
    // | class 1 << 16, word g * 2 + h) of the instances the epilogue's first
    // pass marks, in order
    v4i stash_mem[16], marked_mem[32], keep[KEEP > 0 ? KEEP : 1];
    __attribute__((address_space(5))) v4i *stash = (__attribute__((address_space(5))) v4i *)stash_mem;
    __attribute__((address_space(5))) v4i *marked = (__attribute__((address_space(5))) v4i *)marked_mem;
    asm volatile("" : "+v"(stash), "+v"(marked));
    const int r32 = lane & 31;
#pragma unroll
    for (int cls = 0; cls < 2; ++cls) {
#pragma unroll
        for (int ab = 0; ab < 2; ++ab)
#pragma unroll
            for (int g = 0; g < 2; ++g)
#pragma unroll
                for (int h = 0; h < 2; ++h) acc[ab][g][h] = (v16f){0};
        const int iters = D.quads[cls];
        const uint4 *img = F4.base + F4.class_off[cls];   // quad q at + q * 8 * np
        auto issue = [&](int q) __attribute__((always_inline)) {
            const __amdgpu_buffer_rsrc_t rq =
                __builtin_amdgcn_make_buffer_rsrc((void *)(img + (int64_t)q * 8 * np), (short)0, 128 * np, 0x00020000);
            int8_t *s = lds[q % R];
#pragma unroll
            for (int t = 0; t < kPerWave; ++t) dma16(rq, vo[t], s + (kPerWave * w + t) * 1024);
        };
#pragma unroll
        for (int q = 0; q < R - 1; ++q)
            if (q < iters) issue(q);
        auto stage = [&](int it, bool steady) __attribute__((always_inline)) {
            // kPerWave loads per stage; stage it is complete once at most
            // those of the later stages issued are pending, the barrier
            // publishes every wave's part
            static_assert(kPerWave * (R - 2) < 16, "ring depth");
            if (steady) wait_vmcnt<kPerWave * (R - 2)>();
            else wait_stages<R - 2, kPerWave>(iters - 1 - it);
            stage_barrier();
            __builtin_amdgcn_s_setprio(1);
            const int8_t *sl = lds[it % R];
            // substep s (64 samples): the b fragments of genotypes 0 and 1
            // and the a fragments (g, a block); reads of substep 1 issue
            // between the products of substep 0 (12 reads in a row overfill
            // the LDS command queue)
            v4i fa[2][2][2], fb[2][2];   // [s][g][a block], [s][h]
            auto rd = [&](int s, int row0) __attribute__((always_inline)) {
                return *(const v4i *)(sl + (row0 + r32) * 64 + (((2 * s + (lane >> 5)) ^ m2_swz(r32)) << 4));
            };
#pragma unroll
            for (int h = 0; h < 2; ++h) fb[0][h] = rd(0, 128 * AT + h * 128 + 32 * wb);
#pragma unroll
            for (int g = 0; g < 2; ++g)
#pragma unroll
                for (int ab = 0; ab < 2; ++ab) fa[0][g][ab] = rd(0, g * 64 * AT + 64 * wa + 32 * ab);
            if (steady || it + R - 1 < iters) issue(it + R - 1);
            __builtin_amdgcn_sched_barrier(0);
            auto product = [&](int s, int ab, int g, int h) __attribute__((always_inline)) {
                const v8i a8 = {fa[s][g][ab][0], fa[s][g][ab][1], fa[s][g][ab][2], fa[s][g][ab][3], 0, 0, 0, 0};
                const v8i b8 = {fb[s][h][0], fb[s][h][1], fb[s][h][2], fb[s][h][3], 0, 0, 0, 0};
                // formats 4/4 = e2m1; scale operands 0 select the unscaled product
                acc[ab][g][h] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a8, b8, acc[ab][g][h], 4, 4, 0, 0, 0, 0);
            };
            // substep 0's 8 products, each pair with one read of substep 1
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int g = k >> 1, ab = k & 1;
                if (k < 2) fb[1][k] = rd(1, 128 * AT + k * 128 + 32 * wb);
                fa[1][k >> 1][k & 1] = rd(1, (k >> 1) * 64 * AT + 64 * wa + 32 * (k & 1));
                product(0, ab, g, 0);
                product(0, ab, g, 1);
                __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                product(1, k & 1, k >> 1, 0);
                product(1, k & 1, k >> 1, 1);
            }
            __builtin_amdgcn_s_setprio(0);
        };
        int it = 0;
        for (; it + R - 1 < iters; ++it) stage(it, true);
        for (; it < iters; ++it) stage(it, false);
        // the next class's first loads reuse the ring
        stage_barrier();
#pragma unroll
        for (int ab = 0; ab < 2; ++ab)
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int half = 0; half < 2; ++half)
                    if (cls == 0) {
                        const int i = 2 * (4 * ab + r) + half;
                        if (i < KEEP) keep[i] = pack_counts2(acc, ab, r, half);
                        else stash[i] = pack_counts2(acc, ab, r, half);
                    }
    }
    // epilogue: lane holds column b = b0 + 32 wb + lane % 32 and, per a block
    // ab, rows a = a0 + 64 wa + 32 ab + 8 r + 4 (lane / 32) + q, r, q < 4.  First
    // pass: the bound of each of the 32 instances (ab, r, q), the passing
    // ones marked in `pend` (bit 16 ab + 4 r + q) and their counts kept in
    // `marked`; second pass: one marked instance per lane per round
    const float tau = st->tau;
    const uint64_t tau_combo = st->tau_combo;
    const int b = b0 + 32 * wb + r32;
    const bool b_ok = b < D.n;
    uint32_t sb[2][2] = {{0, 0}, {0, 0}};
    if (b_ok) {
        const uint32_t *sp = D.singles + (size_t)b * 6;
        sb[0][0] = sp[0], sb[0][1] = sp[1], sb[1][0] = sp[3], sb[1][1] = sp[4];
    }
    // the 9 cells of pair (a, b) per class from the counts cnt(cls, g, h)
    auto cells = [&](int a, auto &&cnt, float (&m0)[9], float (&m1)[9]) __attribute__((always_inline)) {
        const int64_t rel = (int64_t)a - P.offset;
        const bool valid = b_ok && a < b && rel >= 0 && (P.step == 1 || rel % P.step == 0);
        const uint32_t *sa = D.singles + (size_t)a * 6;
#pragma unroll
        for (int cls = 0; cls < 2; ++cls) {
            float *m = cls ? m1 : m0;
            m[0] = cnt(cls, 0, 0);
            m[1] = cnt(cls, 0, 1);
            m[3] = cnt(cls, 1, 0);
            m[4] = cnt(cls, 1, 1);
            m[2] = (float)sa[cls * 3 + 0];
            m[5] = (float)sa[cls * 3 + 1];
            m[6] = (float)sb[cls][0];
            m[7] = (float)sb[cls][1];
            m[8] = (float)D.total[cls];
            mobius2(m);
        }
        return valid;
    };
    uint32_t pend = 0;
    int nmarked = 0;
    static_for<8>([&](auto ABR) __attribute__((always_inline)) {
        constexpr int ab = decltype(ABR)::value >> 2, r = decltype(ABR)::value & 3;
        constexpr int i0 = 2 * (4 * ab + r);
        const v4i lo0 = i0 < KEEP ? keep[i0] : stash[i0], lo1 = i0 + 1 < KEEP ? keep[i0 + 1] : stash[i0 + 1];
        static_for<4>([&](auto Q) __attribute__((always_inline)) {
            constexpr int q = decltype(Q)::value;
            float m0[9], m1[9];
            const bool valid = cells(
                a0 + 64 * wa + 32 * ab + 8 * r + 4 * (lane >> 5) + q,
                [&](int cls, int g, int h) __attribute__((always_inline)) {
                    if (cls) return acc[ab][g][h][r * 4 + q];
                    const v4i &lo = g ? lo1 : lo0;
                    return (float)((uint32_t)lo[h * 2 + (q >> 1)] >> (16 * (q & 1)) & 0xffffu);
                },
                m0, m1);
            if (valid && bound_ok<9>(m0, m1, tau, P)) {
                pend |= 1u << (16 * ab + 4 * r + q);
                int wd[4];
#pragma unroll
                for (int g = 0; g < 2; ++g)
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const v4i &lo = g ? lo1 : lo0;
                        wd[g * 2 + h] = (int)((uint32_t)lo[h * 2 + (q >> 1)] >> (16 * (q & 1)) & 0xffffu |
                                              (uint32_t)acc[ab][g][h][r * 4 + q] << 16);
                    }
                marked[nmarked++] = (v4i){wd[0], wd[1], wd[2], wd[3]};
            }
        });
    });
    int taken = 0;
    while (__ballot(pend != 0)) {
        bool want = false;
        float mf = 0.0f;
        uint64_t combo = ~0ull;
        if (pend) {
            const int i = __builtin_ctz(pend), ab = i >> 4, r = (i >> 2) & 3, q = i & 3;
            pend &= pend - 1;
            const int a = a0 + 64 * wa + 32 * ab + 8 * r + 4 * (lane >> 5) + q;
            const v4i wd = marked[taken++];
            float m0[9], m1[9];
            cells(
                a,
                [&](int cls, int g, int h) __attribute__((always_inline)) {
                    const uint32_t v = (uint32_t)wd[g * 2 + h];
                    return (float)(cls ? v >> 16 : v & 0xffffu);
                },
                m0, m1);
            screen_exact<9>(m0, m1, [&] { return prefix_key(P.binom, D.n, &a, 1) + (uint64_t)b; }, tau, tau_combo, P,
                            want, mf, combo);
        }
        push_candidates(want, mf, combo, P, st);
    }
}

__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_mfma2w(const DevData D, const ScanParams P, TopState *st, const Fp4Image F4, int64_t s0, int64_t s1, int T)
{
    __shared__ __attribute__((aligned(1024))) int8_t lds[kM2wRing][kM2wStage];
    // block order: XCD v % 8 takes super-blocks s0 + 8 i + v % 8, 32 blocks each
    const int64_t k = blockIdx.x >> 3, su = s0 + (k >> 5) * 8 + (blockIdx.x & 7);
    if (su >= s1) return;
    mfma2w_tile<kM2wKeep>(D, P, st, F4, su, (int)(k & 31), T, lds);
}

// The 128 x 128 tile: 8 waves, 16 workgroups per super-block
__global__ void __launch_bounds__(2 * kBlock) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_mfma2x(const DevData D, const ScanParams P, TopState *st, const Fp4Image F4, int64_t s0, int64_t s1, int T)
{
    __shared__ __attribute__((aligned(1024))) int8_t lds[kM2xRing][m2w_stage<2>()];
    // block order: XCD v % 8 takes super-blocks s0 + 8 i + v % 8, 16 blocks each
    const int64_t k = blockIdx.x >> 3, su = s0 + (k >> 4) * 8 + (blockIdx.x & 7);
    if (su >= s1) return;
    mfma2w_tile<kM2wKeep, 2, kM2xRing>(D, P, st, F4, su, (int)(k & 15), T, lds);
}

// ---------------------------------------------------------------------------
// Order 3 on the matrix cores.  For a fixed first SNP a, the eight counts
// f(x, y, z), x, y, z in {0, 1}, of the triples (a, b, c) are sums over the
// samples of I_x[a] I_y[b] I_z[c] = I_y[b] . (I_z[c] AND I_x[a]): an fp4
// product of b rows with c rows masked by a row of a.  The other 19 cells
// follow from the pair table, the singles and the class total exactly as in
// the popcount scan (same inclusion-exclusion, same MI and top-N code).
// One workgroup: first SNP a, 64 b's x 32 c's, 4 waves of 32 b x 16 c; each
// v_mfma_f32_16x16x128_f8f6f4 multiplies 16 b rows by 16 c rows over 128
// samples (one bit-table quad).  Work items: the first SNPs a0 .. a0+na-1
// times the tiles (tb <= tc) of the 64-SNP triangle from a's tile on, each
// split into two 32-c halves.
// ---------------------------------------------------------------------------
typedef float v4f __attribute__((ext_vector_type(4)));

// A stage (one 128-sample quad of one class) in LDS: 13 groups of 16 rows x
// 64 bytes -- groups 0..7 the b rows (y, b), 8..11 the c rows (z, c), 12 the
// two a rows (x).  Segment s (16 bytes, 32 samples) of row r of a group is
// stored at segment s ^ m3_swz(r), so that a fragment read (lane l: row
// l % 16, segment l / 16) is free of bank conflicts: measured on gfx950, a
// ds_read_b128 of this pattern takes 4.7 cycles per wave per CU, as a linear
// 1 KB read does, against 8.5-8.8 unswizzled or with (r >> 2) & 3 or
// ((r >> 1) ^ (r >> 3)) & 3.
constexpr int kM3Group = 16 * 64;
constexpr int kM3Stage = 13 * kM3Group;
constexpr int kM3Ring = 4;   // default stages resident: the one read + up to 3 loading
__device__ __forceinline__ int m3_swz(int r) { return (r >> 1) & 3; }

// pair_index for 0 <= x < y < n: one 32 x 32 -> 64-bit product, no signed
// division.
__device__ __forceinline__ uint64_t pair_index_u(uint32_t x, uint32_t y, uint32_t n)
{
    return ((uint64_t)x * (2u * n - x - 1u) >> 1) + (y - x - 1u);
}

// Values of the order-3 matrix-core epilogue that depend only on the first
// SNP a and the lane's last SNP c: c's singles, the pair (a, c) and the rank
// of the prefix (a, a + 1).
struct M3Lane {
    int c;
    bool c_ok;
    // as FP32 (exact): singles of a and c (class, genotype 0 / 1), the pair
    // (a, c) (class, x, z), the class totals
    float sa[2][2], sc[2][2], fac[8], tot[2];
    int64_t rank_a;
};

__device__ __forceinline__ M3Lane m3_lane(const DevData &D, const ScanParams &P, int a, int c)
{
    M3Lane L;
    const int n = D.n;
    L.c = c;
    L.c_ok = c < n && c > a;
    const uint32_t *sa = D.singles + (size_t)a * 6;
    // an out-of-range c reads c = a + 1's rows (its triples are dropped)
    const uint32_t c1 = L.c_ok ? c : a + 1;
    const uint32_t *sp = D.singles + (size_t)c1 * 6;
    const uint4 *pac = (const uint4 *)(P.pair + pair_index_u(a, c1, n) * 8);
    const uint4 u0 = pac[0], u1 = pac[1];
#pragma unroll
    for (int cls = 0; cls < 2; ++cls) {
#pragma unroll
        for (int g = 0; g < 2; ++g) {
            L.sa[cls][g] = (float)sa[cls * 3 + g];
            L.sc[cls][g] = (float)sp[cls * 3 + g];
        }
        L.tot[cls] = (float)D.total[cls];
    }
    const uint32_t f[8] = {u0.x, u0.y, u0.z, u0.w, u1.x, u1.y, u1.z, u1.w};
#pragma unroll
    for (int e = 0; e < 8; ++e) L.fac[e] = (float)f[e];
    L.rank_a = (int64_t)a * (2 * (int64_t)n - a - 1) / 2;   // rank of the prefix (a, a+1)
    return L;
}

// Triple (a, b, L.c): the 8 matrix-core counts cnt(cls, x, y, z), x, y, z in
// {0, 1}, with the pair table, the singles and the class totals give the 27
// cells m0 / m1 of each class (inclusion-exclusion in FP32: the counts are
// exact integers below 2^24, and so is every step); returns whether the
// triple is one of the scan's.  m3_triple: then the shared epilogue.
template <class C>
__device__ __forceinline__ bool m3_cells(const DevData &D, const ScanParams &P, const M3Lane &L, int a, int b, C &&cnt,
                                         float (&m0)[27], float (&m1)[27])
{
    const int n = D.n, c = L.c;
    const bool in_range = L.c_ok && b > a && c > b;   // a < b < c < n
    bool valid = in_range;
    if (P.step != 1 || P.offset != 0) {   // a share of the prefixes (wave-uniform)
        const int64_t rel = L.rank_a + (b - a - 1) - P.offset;
        valid = valid && rel >= 0 && (P.step == 1 || rel % P.step == 0);
    }
    // a triple outside the range reads the rows of (a, a + 1, a + 2) instead
    // (every tile has a < n - 2) and is dropped by `valid`: no select per value
    const uint32_t b1 = in_range ? b : a + 1, c1 = in_range ? c : a + 2;
    uint32_t fab[8], fbc[8], sb[6];
    {
        const uint4 *pab = (const uint4 *)(P.pair + pair_index_u(a, b1, n) * 8);
        const uint4 *pbc = (const uint4 *)(P.pair + pair_index_u(b1, c1, n) * 8);
        const uint4 p0 = pab[0], p1 = pab[1], q0 = pbc[0], q1 = pbc[1];
        fab[0] = p0.x, fab[1] = p0.y, fab[2] = p0.z, fab[3] = p0.w;
        fab[4] = p1.x, fab[5] = p1.y, fab[6] = p1.z, fab[7] = p1.w;
        fbc[0] = q0.x, fbc[1] = q0.y, fbc[2] = q0.z, fbc[3] = q0.w;
        fbc[4] = q1.x, fbc[5] = q1.y, fbc[6] = q1.z, fbc[7] = q1.w;
        const uint32_t *sp = D.singles + (size_t)b1 * 6;
#pragma unroll
        for (int e = 0; e < 6; ++e) sb[e] = sp[e];
    }
#pragma unroll
    for (int cls = 0; cls < 2; ++cls) {
        float *mm = cls ? m1 : m0;
        // mm[x*9 + y*3 + z], x, y, z in {0, 1, 2 = any}
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y) {
#pragma unroll
                for (int z = 0; z < 2; ++z) mm[x * 9 + y * 3 + z] = cnt(cls, x, y, z);
                mm[x * 9 + y * 3 + 2] = (float)fab[cls * 4 + x * 2 + y];
            }
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int z = 0; z < 2; ++z) mm[x * 9 + 6 + z] = L.fac[cls * 4 + x * 2 + z];
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int z = 0; z < 2; ++z) mm[18 + y * 3 + z] = (float)fbc[cls * 4 + y * 2 + z];
        mm[0 * 9 + 8] = L.sa[cls][0];
        mm[1 * 9 + 8] = L.sa[cls][1];
        mm[18 + 0 * 3 + 2] = (float)sb[cls * 3 + 0];
        mm[18 + 1 * 3 + 2] = (float)sb[cls * 3 + 1];
        mm[18 + 6 + 0] = L.sc[cls][0];
        mm[18 + 6 + 1] = L.sc[cls][1];
        mm[26] = L.tot[cls];
        mobius3(mm);
    }
    return valid;
}

__device__ __forceinline__ uint64_t m3_key(const ScanParams &P, int n, int a, int b, int c)
{
    const int pre[2] = {a, b};
    return prefix_key(P.binom, n, pre, 2) + (uint64_t)c;
}

template <class C>
__device__ __forceinline__ void m3_triple(const DevData &D, const ScanParams &P, TopState *st, const M3Lane &L, int a,
                                          int b, float tau, uint64_t tau_combo, C &&cnt)
{
    float m0[27], m1[27];
    const bool valid = m3_cells(D, P, L, a, b, cnt, m0, m1);
    score_and_push<27>(valid, m0, m1, [&] { return m3_key(P, D.n, a, b, L.c); }, tau, tau_combo, P, st);
}

// One workgroup's work item: first SNP a, tile u (see k_mfma3); returns at
// once for items outside the triangle.  Every wave of the workgroup calls it
// with the same arguments.  R: stages of the LDS ring (R - 1 in flight).
template <int R>
__device__ __forceinline__ void mfma3_tile(const DevData &D, const ScanParams &P, TopState *st, const Fp4Image &F,
                                           int a, int64_t u, int T, int8_t (*lds)[kM3Stage])
{
    if (a >= D.n - 2) return;
    const int ta = (a + 1) >> 6, m = T - ta;
    if (u >= (int64_t)m * (m + 1)) return;
    int r, cc;
    tri_tile(u >> 1, m, r, cc);
    const int b0 = (ta + r) * 64, c0 = (ta + cc) * 64 + (int)(u & 1) * 32;
    if (c0 >= D.n) return;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6), wb = w >> 1, wc = w & 1;
    const int np = D.n_pad;
    // staging: each wave issues 4 direct-to-LDS loads per stage, groups 2w
    // and 2w + 1 (b rows (y, b) = (g / 4, b0 + 16 (g % 4) + r)), group 8 + w
    // (c rows (z, c) = (w / 2, c0 + 16 (w % 2) + r)) and lanes 2w, 2w + 1 of
    // group 12 (a row x = r).  Lane l fills row r = l / 4, stored segment
    // l % 4, so it loads the row's segment (l % 4) ^ m3_swz(r).  Byte
    // offsets within a quad of the image (< 128 n_pad).
    const int lr = lane >> 2, ls = (lane & 3) ^ m3_swz(lr);
    const uint32_t ub0 = 16u * (uint32_t)(4 * ((w >> 1) * np + b0 + 32 * (w & 1) + lr) + ls);        // group 2w
    const uint32_t ub1 = 16u * (uint32_t)(4 * ((w >> 1) * np + b0 + 32 * (w & 1) + 16 + lr) + ls);   // 2w + 1
    const uint32_t uc = 16u * (uint32_t)(4 * ((w >> 1) * np + c0 + 16 * (w & 1) + lr) + ls);         // 8 + w
    const bool a_lane = (lane >> 1) == w;                                       // lanes 2w, 2w + 1 (< 8)
    const uint32_t ua = 16u * (uint32_t)(4 * ((lr & 1) * np + a) + ls);        // group 12, row x = lr
    v4f acc[2][2][2][2][2];   // [class][x][y][b block][z]
    // fragment reads: lane reads row rr of a group at segment kc (32 samples
    // 32 kc ..), stored at segment kc ^ m3_swz(rr); the a rows (m3_swz = 0)
    // are read by every lane of a 16-lane group at once (broadcast)
    const int rr = lane & 15, kc = lane >> 4;
    const int fo = rr * 64 + ((kc ^ m3_swz(rr)) << 4), ao = kc << 4;
#pragma unroll
    for (int cls = 0; cls < 2; ++cls) {
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y)
#pragma unroll
                for (int bb = 0; bb < 2; ++bb)
#pragma unroll
                    for (int z = 0; z < 2; ++z) acc[cls][x][y][bb][z] = (v4f){0.f, 0.f, 0.f, 0.f};
        const int iters = D.quads[cls];
        const uint4 *img = F.base + F.class_off[cls];
        const int64_t qstride = 8 * (int64_t)np;   // uint4 per quad: 2 genotypes x n_pad x 4
        // resource over quad q of the class image (8 n_pad uint4)
        auto quad_rsrc3 = [&](int q) __attribute__((always_inline)) {
            return __builtin_amdgcn_make_buffer_rsrc((void *)(img + q * qstride), (short)0, 128 * np, 0x00020000);
        };
        // the b and c rows, and (two lanes of each wave) the a rows, of quad
        // q: 4 loads per wave in this order
        auto issue_bc = [&](int q) __attribute__((always_inline)) {
            const __amdgpu_buffer_rsrc_t rq = quad_rsrc3(q);
            int8_t *s = lds[q % R];
            dma16(rq, ub0, s + (2 * w) * kM3Group);
            dma16(rq, ub1, s + (2 * w + 1) * kM3Group);
            dma16(rq, uc, s + (8 + w) * kM3Group);
        };
        auto issue_a = [&](int q) __attribute__((always_inline)) {
            if (a_lane) dma16(quad_rsrc3(q), ua, lds[q % R] + 12 * kM3Group);
        };
        auto issue = [&](int q) __attribute__((always_inline)) {
            issue_bc(q);
            issue_a(q);
        };
        struct Frag {
            v4i ma[2], fb[2][2], fc[2];
        };
        // the a and c fragments first: the masks need them, the products
        // need the b fragments only after that
        auto read = [&](Frag &f, int it) __attribute__((always_inline)) {
            const int8_t *sb = lds[it % R];
#pragma unroll
            for (int x = 0; x < 2; ++x) f.ma[x] = *(const v4i *)(sb + 12 * kM3Group + x * 64 + ao);
#pragma unroll
            for (int z = 0; z < 2; ++z) f.fc[z] = *(const v4i *)(sb + (8 + 2 * z + wc) * kM3Group + fo);
#pragma unroll
            for (int y = 0; y < 2; ++y)
#pragma unroll
                for (int bb = 0; bb < 2; ++bb)
                    f.fb[y][bb] = *(const v4i *)(sb + (4 * y + 2 * wb + bb) * kM3Group + fo);
        };
        // the a mask goes on the narrower operand (a wave's 16 c rows per
        // z, against 32 b rows per y): 4 masked fragments per stage
        auto multiply = [&](const Frag &f) __attribute__((always_inline)) {
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int z = 0; z < 2; ++z) {
                    const v4i mc = f.fc[z] & f.ma[x];   // nibble 0x2 & 0x2 = 1.0, else 0
                    const v8i c8 = {mc[0], mc[1], mc[2], mc[3], 0, 0, 0, 0};
#pragma unroll
                    for (int y = 0; y < 2; ++y)
#pragma unroll
                        for (int bb = 0; bb < 2; ++bb) {
                            const v8i b8 = {f.fb[y][bb][0], f.fb[y][bb][1], f.fb[y][bb][2], f.fb[y][bb][3],
                                            0, 0, 0, 0};
                            acc[cls][x][y][bb][z] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(
                                b8, c8, acc[cls][x][y][bb][z], 4, 4, 0, 0, 0, 0);
                        }
                }
        };
#pragma unroll
        for (int q = 0; q < R - 1; ++q)
            if (q < iters) issue(q);
        // the stages with at least R - 1 stages after them run a body without
        // branches (constant wait, always one stage to issue), so that the
        // fragment reads, masks and products share a basic block and each
        // product waits only for its own operands; the last R - 1 stages run
        // the general body
        auto stage = [&](int it, bool steady) __attribute__((always_inline)) {
            // this wave's loads of stage it are complete once at most
            // those of the stages after it (min(iters - 1 - it, R - 2),
            // 4 loads each) are pending; the barrier then publishes every
            // wave's part
            static_assert(R >= 2 && 4 * (R - 2) < 16, "ring depth");
            if (steady) wait_vmcnt<4 * (R - 2)>();
            else wait_stages<R - 2, 4>(iters - 1 - it);
            stage_barrier();
            // the stage (fragment reads, products) issues ahead of the other
            // wave's epilogue
            __builtin_amdgcn_s_setprio(1);
            Frag f;
            read(f, it);
            // every wave is past stage it - 1, whose slot now takes stage it +
            // R - 1; the a rows' load (two lanes, a branch) after the products
            const bool more = steady || it + R - 1 < iters;
            if (more) issue_bc(it + R - 1);
            multiply(f);
            if (steady) {
                // issue order: the 8 fragment reads, the first mask, then each
                // remaining mask word right behind a product (it issues while
                // the product occupies the matrix core), the last 4 products
                // (3 % faster than the compiler's order, which put all 16
                // masks ahead of the products)
                __builtin_amdgcn_sched_group_barrier(0x100, 8, 0);
                __builtin_amdgcn_sched_group_barrier(0x002, 4, 0);
#pragma unroll
                for (int i = 0; i < 12; ++i) {
                    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                    __builtin_amdgcn_sched_group_barrier(0x002, 1, 0);
                }
                __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);
            }
            __builtin_amdgcn_s_setprio(0);
            if (more) issue_a(it + R - 1);
        };
        int it = 0;
        for (; it + R - 1 < iters; ++it) stage(it, true);
        for (; it < iters; ++it) stage(it, false);
        // the next class's first loads reuse the ring: every wave must be
        // done reading this class's last stages
        stage_barrier();
    }
    // epilogue: lane holds column c = c0 + wc*16 + lane%16 and, per b block
    // bb, rows b = b0 + wb*32 + 16 bb + 4(lane/16) + j, j < 4
    const float tau = st->tau;
    const uint64_t tau_combo = st->tau_combo;
    const M3Lane L = m3_lane(D, P, a, c0 + wc * 16 + rr);
    // eight instances (bb, j), so that every accumulator index is a constant
    static_for<8>([&](auto BJ) __attribute__((always_inline)) {
        constexpr int bb = decltype(BJ)::value >> 2, j = decltype(BJ)::value & 3;
        m3_triple(D, P, st, L, a, b0 + wb * 32 + bb * 16 + 4 * (lane >> 4) + j, tau, tau_combo,
                  [&](int cls, int x, int y, int z) __attribute__((always_inline)) {
                      return acc[cls][x][y][bb][z][j];
                  });
    });
}

template <int R>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2, 2)))   // 256 registers, no spills
k_mfma3(const DevData D, const ScanParams P, TopState *st, const Fp4Image F, int a0, int na, int T)
{
    __shared__ __attribute__((aligned(1024))) int8_t lds[R][kM3Stage];
    // block order: workgroup L runs on XCD L % 8; XCD k takes the tiles
    // u = 8 i + k, i = 0, 1, ..., and for each tile all na first SNPs in a
    // row, so that the workgroups resident on one XCD share their b and c
    // rows in its L2 (only the a rows differ)
    const int64_t L = blockIdx.x, s = L >> 3;
    mfma3_tile<R>(D, P, st, F, a0 + (int)(s % na), (s / na) * 8 + (L & 7), T, lds);
}

// Wide tile: one workgroup takes first SNP a and a whole 64 x 64 tile of
// (b, c); wave w multiplies all 64 b rows by the 16 c rows c0 + 16 w ...: 32
// products per 128-sample stage against 12 fragment reads and 16 masks (16,
// 8 and 16 per 16 products in mfma3_tile), and 16 KB staged per 128 products
// (12 KB per 64).  One class's counts fill the 128 accumulators; the first
// class's wait as 16-bit pairs (64 registers) while the second runs, so
// every class must hold fewer than 2^16 samples.  Stage layout: groups 0..7
// the b rows (y, 16-row block bb), 8..15 the c rows (z, block), 16 the a
// rows, each stored as in mfma3_tile.
constexpr int kM3wStage = 17 * kM3Group;
// Words of the first class's 16 packed count words that stay in registers
// through the second class's stages (the rest wait in private memory, whose
// lines the operand stream evicts from L2).
constexpr int kM3wKeep = 4;

// Counts of b block bb = i / 4, (x, y) = (i / 2 % 2, i % 2) as 16-bit pairs:
// word z * 2 + j / 2 holds f(x, y, z) of rows 4 (lane / 16) + j, in the low
// (j even) or high half (every count is below 2^16).
__device__ __forceinline__ v4i pack_counts(const v4f (&acc)[2][2][4][2], int i)
{
    auto pk = [](float lo, float hi) __attribute__((always_inline)) {
        return (int)((uint32_t)lo | (uint32_t)hi << 16);
    };
    const int bb = i >> 2, x = (i >> 1) & 1, y = i & 1;
    const v4f &v0 = acc[x][y][bb][0], &v1 = acc[x][y][bb][1];
    return (v4i){pk(v0[0], v0[1]), pk(v0[2], v0[3]), pk(v1[0], v1[1]), pk(v1[2], v1[3])};
}

template <int R>
__device__ __forceinline__ void mfma3w_tile(const DevData &D, const ScanParams &P, TopState *st, const Fp4Image &F,
                                            int a, int64_t u, int T, int8_t (*lds)[kM3wStage])
{
    if (a >= D.n - 2) return;
    const int ta = (a + 1) >> 6, m = T - ta;
    if (u >= (int64_t)m * (m + 1) / 2) return;
    int r, cc;
    tri_tile(u, m, r, cc);
    const int b0 = (ta + r) * 64, c0 = (ta + cc) * 64;
    if (c0 >= D.n) return;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int np = D.n_pad;
    // staging: wave w loads b groups 2w, 2w + 1 (y = w / 2, blocks 2 (w % 2)
    // and the next; +1,024 bytes), its own c rows (groups 8 + w and 12 + w:
    // z = 0, 1, SNPs c0 + 16 w ..) and, lanes 0 .. 7, its own copy of the a
    // rows (bytes 128 w .. of group 16).  The c and a rows a wave reads are
    // its own loads, complete after its own vmcnt wait: the next stage's c
    // and a fragments and masks are read behind the current stage's x = 1
    // products, before the barrier that publishes the shared b rows
    const int lr = lane >> 2, ls = (lane & 3) ^ m3_swz(lr);
    const uint32_t ub = 16u * (uint32_t)(4 * ((w >> 1) * np + b0 + 32 * (w & 1) + lr) + ls);
    const uint32_t uc0 = 16u * (uint32_t)(4 * (c0 + 16 * w + lr) + ls), uc1 = uc0 + 64u * (uint32_t)np;
    const bool a_lane = lane < 8;
    const uint32_t ua = 16u * (uint32_t)(4 * ((lr & 1) * np + a) + ls);   // lanes 0 .. 7: row x = lr
    v4f acc[2][2][4][2];   // [x][y][b block][z], the class being scanned
    // the first class's counts (pack_counts, 16 words) in private memory,
    // out of the registers the second class needs; `marked`: both classes'
    // counts of the instances the epilogue's first pass marks.  The pointers
    // pass through an empty asm statement, so the compiler keeps plain
    // 16-byte private stores and loads (a volatile access would be uncached
    // and waited for one by one)
    v4i stash_mem[16], marked_mem[32], keep[kM3wKeep];
    __attribute__((address_space(5))) v4i *stash = (__attribute__((address_space(5))) v4i *)stash_mem;
    __attribute__((address_space(5))) v4i *marked = (__attribute__((address_space(5))) v4i *)marked_mem;
    asm volatile("" : "+v"(stash), "+v"(marked));
    const int rr = lane & 15, kc = lane >> 4;
    const int fo = rr * 64 + ((kc ^ m3_swz(rr)) << 4), ao = kc << 4;
#pragma unroll
    for (int cls = 0; cls < 2; ++cls) {
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y)
#pragma unroll
                for (int bb = 0; bb < 4; ++bb)
#pragma unroll
                    for (int z = 0; z < 2; ++z) acc[x][y][bb][z] = (v4f){0.f, 0.f, 0.f, 0.f};
        const int iters = D.quads[cls];
        const uint4 *img = F.base + F.class_off[cls];
        const int64_t qstride = 8 * (int64_t)np;
        auto quad_rsrc = [&](int q) __attribute__((always_inline)) {
            return __builtin_amdgcn_make_buffer_rsrc((void *)(img + q * qstride), (short)0, 128 * np, 0x00020000);
        };
        auto issue_bc = [&](int q) __attribute__((always_inline)) {
            const __amdgpu_buffer_rsrc_t rq = quad_rsrc(q);
            int8_t *s = lds[q % R];
            dma16(rq, ub, s + (2 * w) * kM3Group);
            dma16(rq, ub + 1024u, s + (2 * w + 1) * kM3Group);
            dma16(rq, uc0, s + (8 + w) * kM3Group);
            dma16(rq, uc1, s + (12 + w) * kM3Group);
        };
        auto issue_a = [&](int q) __attribute__((always_inline)) {
            if (a_lane) dma16(quad_rsrc(q), ua, lds[q % R] + 16 * kM3Group + 128 * w);
        };
#pragma unroll
        for (int q = 0; q < R - 1; ++q)
            if (q < iters) {
                issue_bc(q);
                issue_a(q);
            }
        // 5 loads per stage (4 issued early in a stage, the a rows' at its
        // end; the a load counts for every lane)
        static_assert(R == 3 || R == 4, "the waits below count rings of 3 or 4 stages");
        // the masks of stage q (x, z): this wave's c rows AND its a rows
        auto masks = [&](int q, v4i (&mk)[2][2]) __attribute__((always_inline)) {
            const int8_t *sq = lds[q % R];
            v4i ma[2], fc[2];
#pragma unroll
            for (int x = 0; x < 2; ++x) ma[x] = *(const v4i *)(sq + 16 * kM3Group + 128 * w + x * 64 + ao);
#pragma unroll
            for (int z = 0; z < 2; ++z) fc[z] = *(const v4i *)(sq + (8 + 4 * z + w) * kM3Group + fo);
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int z = 0; z < 2; ++z) mk[x][z] = fc[z] & ma[x];   // nibble 0x2 & 0x2 = 1.0, else 0
        };
        // stage 0's masks: its loads complete once at most those of stages
        // 1 .. min(R - 1, iters) - 1 are pending
        v4i mA[2][2], mB[2][2];
        wait_stages<R - 2, 5>(iters - 1);
        masks(0, mA);
        auto stage = [&](int it, bool steady, v4i (&mc)[2][2], v4i (&mn)[2][2]) __attribute__((always_inline)) {
            // this wave's loads of stage it are complete once at most those
            // of the stages after it (min(iters - 1 - it, R - 2), 5 loads
            // each) are pending; the barrier then publishes every wave's b rows
            if (steady) wait_vmcnt<5 * (R - 2)>();
            else wait_stages<R - 2, 5>(iters - 1 - it);
            stage_barrier();
            __builtin_amdgcn_s_setprio(1);
            const int8_t *sb = lds[it % R];
            v4i fb[8];
#pragma unroll
            for (int k = 0; k < 2; ++k) fb[k] = *(const v4i *)(sb + k * kM3Group + fo);
            const bool more = steady || it + R - 1 < iters;
            if (more) issue_bc(it + R - 1);
            __builtin_amdgcn_sched_barrier(0);
            auto product = [&](int x, int k, int z) __attribute__((always_inline)) {
                const v8i b8 = {fb[k][0], fb[k][1], fb[k][2], fb[k][3], 0, 0, 0, 0};
                const v8i c8 = {mc[x][z][0], mc[x][z][1], mc[x][z][2], mc[x][z][3], 0, 0, 0, 0};
                acc[x][k >> 2][k & 3][z] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(
                    b8, c8, acc[x][k >> 2][k & 3][z], 4, 4, 0, 0, 0, 0);
            };
            // the next stage's c and a fragments come from this wave's own
            // loads: complete once at most those issued after them are pending
            // (4 stages: stage it + 2's 5 and stage it + 3's 4 early ones; 3
            // stages: stage it + 2's 4 early ones; where those stages exist)
            const bool next = steady || it + 1 < iters;
            const int8_t *sn = lds[(it + 1) % R];
            v4i na[2], nc[2];
            // x = 0: b fragment k (y = k / 4, block k % 4): the read of
            // fragment k + 2 (k < 6) or of the next stage's a and c fragments
            // (k = 6, 7), the two products
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                if (k < 6) {
                    fb[k + 2] = *(const v4i *)(sb + (k + 2) * kM3Group + fo);
                } else if (next) {
                    if (k == 6) {
                        if constexpr (R == 4) {
                            if (steady || it + 3 < iters) wait_vmcnt<9>();
                            else if (it + 2 < iters) wait_vmcnt<5>();
                            else wait_vmcnt<0>();
                        } else {
                            if (steady || it + 2 < iters) wait_vmcnt<4>();
                            else wait_vmcnt<0>();
                        }
                    }
                    const int x = k - 6;
                    na[x] = *(const v4i *)(sn + 16 * kM3Group + 128 * w + x * 64 + ao);
                    nc[x] = *(const v4i *)(sn + (8 + 4 * x + w) * kM3Group + fo);
                }
                product(0, k, 0);
                product(0, k, 1);
                __builtin_amdgcn_sched_barrier(0);
            }
            // x = 1: two products, then two words of the next stage's masks
            // (nibble 0x2 & 0x2 = 1.0, else 0), issued while the products run
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                product(1, k, 0);
                product(1, k, 1);
                if (next) {
#pragma unroll
                    for (int e = 0; e < 2; ++e) {
                        const int i = 2 * k + e, x = i >> 3, z = (i >> 2) & 1, d = i & 3;
                        // the empty asm holds the AND here: left free, the
                        // compiler sinks all 16 behind the last product
                        int v = nc[z][d] & na[x][d];
                        asm volatile("" : "+v"(v));
                        mn[x][z][d] = v;
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
            }
            if (more) issue_a(it + R - 1);
            __builtin_amdgcn_s_setprio(0);
        };
        // stages with at least R - 1 stages after them (and one more, the
        // pair's second) run a body without branches; mask registers
        // alternate between the two stages of a pair
        int it = 0;
        for (; it + R < iters; it += 2) {
            stage(it, true, mA, mB);
            stage(it + 1, true, mB, mA);
        }
        for (; it < iters; it += 2) {
            stage(it, false, mA, mB);
            if (it + 1 < iters) stage(it + 1, false, mB, mA);
        }
        // the next class's first loads reuse the ring: every wave must be
        // done reading this class's last stages
        stage_barrier();
        if (cls == 0) {
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                if (i < kM3wKeep) keep[i] = pack_counts(acc, i);
                else stash[i] = pack_counts(acc, i);
            }
        }
    }
    // epilogue: lane holds column c = c0 + 16 w + lane % 16 and, per b block
    // bb, rows b = b0 + 16 bb + 4 (lane / 16) + j, j < 4
    const float tau = st->tau;
    const uint64_t tau_combo = st->tau_combo;
    const M3Lane L = m3_lane(D, P, a, c0 + 16 * w + rr);
    // first pass: the bound of each of the 16 instances (bb, j); a lane's
    // instances that pass are marked in `pend` (bit 4 bb + j), their counts
    // (f(x, y, z) of class 0 | class 1 << 16, word (x * 2 + y) * 2 + z) kept
    // in `marked` in the same order.  Second pass: each round scores one
    // marked instance per lane, so the screening MI (54 logarithms) runs
    // about as often as the most marked lane has marks, not once per
    // instance in which any lane of the wave passes the bound (most
    // instances, at a top-100 threshold)
    uint32_t pend = 0;
    int nmarked = 0;
    static_for<4>([&](auto BB) __attribute__((always_inline)) {
        constexpr int bb = decltype(BB)::value;
        v4i lo[4];
#pragma unroll
        for (int xy = 0; xy < 4; ++xy) lo[xy] = 4 * bb + xy < kM3wKeep ? keep[4 * bb + xy] : stash[4 * bb + xy];
        static_for<4>([&](auto J) __attribute__((always_inline)) {
            constexpr int j = decltype(J)::value;
            float m0[27], m1[27];
            const bool valid = m3_cells(
                D, P, L, a, b0 + 16 * bb + 4 * (lane >> 4) + j,
                [&](int cls, int x, int y, int z) __attribute__((always_inline)) {
                    return cls ? acc[x][y][bb][z][j]
                               : (float)((uint32_t)lo[x * 2 + y][z * 2 + (j >> 1)] >> (16 * (j & 1)) & 0xffffu);
                },
                m0, m1);
            if (valid && bound_ok<27>(m0, m1, tau, P)) {
                pend |= 1u << (4 * bb + j);
                int wd[8];
#pragma unroll
                for (int x = 0; x < 2; ++x)
#pragma unroll
                    for (int y = 0; y < 2; ++y)
#pragma unroll
                        for (int z = 0; z < 2; ++z)
                            wd[(x * 2 + y) * 2 + z] =
                                (int)((uint32_t)lo[x * 2 + y][z * 2 + (j >> 1)] >> (16 * (j & 1)) & 0xffffu |
                                      (uint32_t)acc[x][y][bb][z][j] << 16);
                marked[2 * nmarked] = (v4i){wd[0], wd[1], wd[2], wd[3]};
                marked[2 * nmarked + 1] = (v4i){wd[4], wd[5], wd[6], wd[7]};
                ++nmarked;
            }
        });
    });
    int taken = 0;
    while (__ballot(pend != 0)) {
        bool want = false;
        float mf = 0.0f;
        uint64_t combo = ~0ull;
        if (pend) {
            const int i = __builtin_ctz(pend), bb = i >> 2, j = i & 3;
            pend &= pend - 1;
            const int b = b0 + 16 * bb + 4 * (lane >> 4) + j;
            const v4i w0 = marked[2 * taken], w1 = marked[2 * taken + 1];
            ++taken;
            float m0[27], m1[27];
            m3_cells(
                D, P, L, a, b,
                [&](int cls, int x, int y, int z) __attribute__((always_inline)) {
                    const int e = (x * 2 + y) * 2 + z;
                    const uint32_t v = (uint32_t)(e < 4 ? w0[e] : w1[e - 4]);
                    return (float)(cls ? v >> 16 : v & 0xffffu);
                },
                m0, m1);
            screen_exact<27>(m0, m1, [&] { return m3_key(P, D.n, a, b, L.c); }, tau, tau_combo, P, want, mf, combo);
        }
        push_candidates(want, mf, combo, P, st);
    }
}

template <int R>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_mfma3w(const DevData D, const ScanParams P, TopState *st, const Fp4Image F, int a0, int na, int T)
{
    __shared__ __attribute__((aligned(1024))) int8_t lds[R][kM3wStage];
    // block order as in k_mfma3 (XCD L % 8 takes tiles u = 8 i + L % 8)
    const int64_t L = blockIdx.x, s = L >> 3;
    mfma3w_tile<R>(D, P, st, F, a0 + (int)(s % na), (s / na) * 8 + (L & 7), T, lds);
}


// ---------------------------------------------------------------------------
// Small helpers of the searches (the ones in a fixed device block use no
// allocation and no memset): the binomial table, the search start, result
// unpacking.
// ---------------------------------------------------------------------------
// binom[j * (n + 1) + x] = C(x, j), saturated at UINT64_MAX, for j <= m.
// C(x - j + i, i) = C(x - j + i - 1, i - 1) (x - j + i) / i is exact at every
// step; with r = q i + rem the product splits as q v + (rem v) / i, where
// rem v is a multiple of i and below 2^35, so no 128-bit division is needed.
// The values grow with i (x >= j), so a saturated step stays saturated.
__global__ void k_binom_table(uint64_t *binom, uint64_t n, int m)
{
    const uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, stride = n + 1;
    if (id >= (uint64_t)(m + 1) * stride) return;
    const int j = (int)(id / stride);
    const uint64_t x = id % stride;
    uint64_t r = 0;
    if ((uint64_t)j <= x) {
        r = 1;
        for (int i = 1; i <= j && r != UINT64_MAX; ++i) {
            const uint64_t v = x - (uint64_t)j + (uint64_t)i, q = r / (uint64_t)i, rem = r % (uint64_t)i;
            const uint64_t lo = q * v, part = rem * v / (uint64_t)i;
            r = (__umul64hi(q, v) != 0 || lo > UINT64_MAX - part) ? UINT64_MAX : lo + part;
        }
    }
    binom[id] = r;
}

// A search's start: the top-N state reset (empty list, threshold -inf)
// and, when `hist` is set, the sampling histogram cleared (a kernel rather
// than copies and fills, so the start is one launch).
__global__ void k_search_begin(TopState *st, uint32_t *hist)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        TopState init{};
        init.tau = -INFINITY;
        init.tau_combo = ~0ull;
        *st = init;
    }
    if (hist)
        for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < kHistBins; b += gridDim.x * blockDim.x) hist[b] = 0;
}

// Top-N entries -> ascending SNP indices (order per entry, the key's
// combination by lexicographic unranking) and float MI, as the host's
// unpack_results does.
__global__ void k_unpack_top(const Cand *top, uint32_t count, int order, int n, const uint64_t *binom, int32_t *combos,
                             float *mi)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    int c[kMaxOrder];
    unrank(top[i].combo, n, order, binom, c);
    for (int q = 0; q < order; ++q) combos[(size_t)i * order + q] = c[q];
    mi[i] = top[i].mi;
}

// ---------------------------------------------------------------------------
// Top-N merge (one workgroup): current list + candidates -> best N, unique.
// ---------------------------------------------------------------------------
// Sorts buf[0, L) (L a power of two, padded with empty entries) by key,
// then writes the first N unique non-empty entries to top; returns how many
// unique entries there were (all threads).
__device__ __forceinline__ uint32_t sort_unique(Cand *buf, uint32_t L, Cand *top, uint32_t N, uint32_t *psum)
{
    const int tid = threadIdx.x;
    for (uint32_t k = 2; k <= L; k <<= 1) {
        for (uint32_t jj = k >> 1; jj > 0; jj >>= 1) {
            for (uint32_t i = tid; i < L; i += 1024) {
                const uint32_t ixj = i ^ jj;
                if (ixj > i) {
                    const Cand a = buf[i], b = buf[ixj];
                    const bool desc = (i & k) == 0;
                    const bool sw = desc ? better(b.mi, b.combo, a.mi, a.combo) : better(a.mi, a.combo, b.mi, b.combo);
                    if (sw) { buf[i] = b; buf[ixj] = a; }
                }
            }
            __syncthreads();
        }
    }
    // unique, non-empty entries in order; contiguous segments per thread
    const uint32_t seg = (L + 1023) / 1024;
    const uint32_t s0 = min(L, tid * seg), s1 = min(L, s0 + seg);
    auto keep = [&](uint32_t i) {
        const bool empty = buf[i].combo == ~0ull && buf[i].mi == -INFINITY;
        const bool dup = i > 0 && buf[i].combo == buf[i - 1].combo && buf[i].mi == buf[i - 1].mi &&
                         buf[i].tag == buf[i - 1].tag;
        return !empty && !dup;
    };
    uint32_t local = 0;
    for (uint32_t i = s0; i < s1; ++i) local += keep(i);
    psum[tid] = local;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const uint32_t v = tid >= off ? psum[tid - off] : 0;
        __syncthreads();
        psum[tid] += v;
        __syncthreads();
    }
    uint32_t pos = psum[tid] - local;
    for (uint32_t i = s0; i < s1; ++i) {
        if (keep(i)) {
            if (pos < N) top[pos] = buf[i];
            ++pos;
        }
    }
    const uint32_t unique = psum[1023];
    __syncthreads();
    return unique;
}

// Body shared by the single-scan and the batched merge (1024 threads).
// Many candidates (a first slice scanned with no threshold pushes every
// combination) are filtered first: a histogram of MI (hist_bin) over the
// current list and the candidates gives the highest bin b with at least 2N
// entries at or above it.  At most N entries are duplicates of others (a
// re-scanned range repeats entries of the list, at most one copy each), so
// at least N unique entries lie in bins >= b and every entry below b ranks
// after them: only the entries in bins >= b are sorted.
__device__ __forceinline__ void merge_topn(TopState *st, Cand *top, const Cand *cand, uint32_t N, uint32_t cap,
                                           Cand *buf, uint32_t *psum)
{
    const int tid = threadIdx.x;
    const uint32_t count = min(st->cand_count, cap);
    uint32_t n_top = st->n_top;
    bool merged = false;
    if (count > (uint32_t)kMergeSlots / 2 && N > 0) {
        static_assert(kHistBins * sizeof(uint32_t) <= kMergeSlots * sizeof(Cand), "histogram in the merge buffer");
        static_assert(kHistBins == 4 * 1024, "four bins per thread");
        uint32_t *hist = (uint32_t *)buf;
        for (int b = tid; b < kHistBins; b += 1024) hist[b] = 0;
        __syncthreads();
        for (uint32_t i = tid; i < count; i += 1024) atomicAdd(&hist[hist_bin(cand[i].mi)], 1u);
        for (uint32_t i = tid; i < n_top; i += 1024) atomicAdd(&hist[hist_bin(top[i].mi)], 1u);
        __syncthreads();
        // suffix sums over the threads' four-bin groups (thread t: bins 4t .. 4t + 3)
        uint32_t grp[4], local = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) local += grp[q] = hist[4 * tid + q];
        psum[tid] = local;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const uint32_t v = tid + off < 1024 ? psum[tid + off] : 0;
            __syncthreads();
            psum[tid] += v;
            __syncthreads();
        }
        const uint32_t need = 2 * N;
        const uint32_t above = tid + 1 < 1024 ? psum[tid + 1] : 0;   // entries in the bins of higher threads
        int sel = -1;                                                   // this thread's pick, if it holds b
        if (psum[tid] >= need && above < need) {
            uint32_t acc = above;
            for (int q = 3; q >= 0 && sel < 0; --q) {
                acc += grp[q];
                if (acc >= need) sel = 4 * tid + q;
            }
        }
        __syncthreads();
        if (tid == 0) psum[0] = 0;   // no pick: keep every bin
        __syncthreads();
        if (sel >= 0) psum[0] = (uint32_t)sel;
        __syncthreads();
        const uint32_t bmin = psum[0];
        __syncthreads();
        // compact the entries in bins >= bmin into buf (order is irrelevant:
        // they are sorted next); counts first, so that the total is known
        uint32_t mine = 0;
        for (uint32_t i = tid; i < count; i += 1024) mine += hist_bin(cand[i].mi) >= bmin;
        for (uint32_t i = tid; i < n_top; i += 1024) mine += hist_bin(top[i].mi) >= bmin;
        psum[tid] = mine;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const uint32_t v = tid >= off ? psum[tid - off] : 0;
            __syncthreads();
            psum[tid] += v;
            __syncthreads();
        }
        const uint32_t total = psum[1023];
        if (total <= (uint32_t)kMergeSlots) {
            uint32_t pos = psum[tid] - mine;
            __syncthreads();   // the histogram (the start of buf) is read no more
            for (uint32_t i = tid; i < n_top; i += 1024)
                if (hist_bin(top[i].mi) >= bmin) buf[pos++] = top[i];
            for (uint32_t i = tid; i < count; i += 1024)
                if (hist_bin(cand[i].mi) >= bmin) buf[pos++] = cand[i];
            uint32_t L = 2;
            while (L < total) L <<= 1;
            for (uint32_t i = total + tid; i < L; i += 1024) {
                Cand e;
                e.mi = -INFINITY;
                e.tag = 0;
                e.combo = ~0ull;
                buf[i] = e;
            }
            __syncthreads();
            n_top = min(N, sort_unique(buf, L, top, N, psum));
            merged = true;
        }
        __syncthreads();
    }
    for (uint32_t done = 0; !merged && done < count;) {
        const uint32_t take = min(count - done, (uint32_t)kMergeSlots - n_top);
        const uint32_t total = n_top + take;
        uint32_t L = 2;
        while (L < total) L <<= 1;
        for (uint32_t i = tid; i < L; i += 1024) {
            Cand e;
            if (i < n_top) e = top[i];
            else if (i < total) e = cand[done + (i - n_top)];
            else { e.mi = -INFINITY; e.tag = 0; e.combo = ~0ull; }
            buf[i] = e;
        }
        __syncthreads();
        n_top = min(N, sort_unique(buf, L, top, N, psum));
        done += take;
    }
    __syncthreads();
    if (tid == 0) {
        st->n_top = n_top;
        // with fewer than N entries the threshold stays as it was (-inf, or
        // the sampled starting threshold, which the final N-th best exceeds)
        if (n_top >= N && N > 0) {
            st->tau = top[N - 1].mi;
            st->tau_combo = top[N - 1].combo;
        }
        st->cand_count = 0;
    }
}

// ---------------------------------------------------------------------------
// Top-N merge for N above kMaxOutputs (global memory).  The host reads the
// list's and the candidates' sizes first (a sync per chunk on this path only)
// and sorts next_pow2(n_top + count) entries: gather, one bitonic step per
// launch, then one workgroup keeps the first N unique entries in order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ Cand empty_cand()
{
    Cand e;
    e.mi = -INFINITY;
    e.tag = 0;
    e.combo = ~0ull;
    return e;
}

__global__ void k_big_gather(const Cand *top, uint32_t n_top, const Cand *cand, uint32_t count, Cand *work, uint32_t L)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L) return;
    work[i] = i < n_top ? top[i] : i < n_top + count ? cand[i - n_top] : empty_cand();
}

// One compare-exchange step (k, jj) of a bitonic sort of L entries into
// key order (MI descending, combination ascending).
__global__ void k_big_bitonic(Cand *work, uint32_t L, uint32_t k, uint32_t jj)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L) return;
    const uint32_t ixj = i ^ jj;
    if (ixj <= i) return;
    const Cand a = work[i], b = work[ixj];
    const bool desc = (i & k) == 0;
    const bool sw = desc ? better(b.mi, b.combo, a.mi, a.combo) : better(a.mi, a.combo, b.mi, b.combo);
    if (sw) {
        work[i] = b;
        work[ixj] = a;
    }
}

// The first N unique non-empty entries of the sorted work buffer -> top;
// state: list size, threshold, candidates consumed, overflow -> re-scan list.
__global__ void __launch_bounds__(1024)
k_big_finish(const Cand *work, uint32_t L, Cand *top, uint32_t N, TopState *st, uint32_t chunk_id, uint32_t *redo)
{
    __shared__ uint32_t psum[1024];
    const int tid = threadIdx.x;
    const uint32_t seg = (L + 1023) / 1024;
    const uint32_t s0 = min(L, tid * seg), s1 = min(L, s0 + seg);
    auto keep = [&](uint32_t i) {
        const Cand c = work[i];
        const bool empty = c.combo == ~0ull && c.mi == -INFINITY;
        const bool dup = i > 0 && c.combo == work[i - 1].combo && c.mi == work[i - 1].mi && c.tag == work[i - 1].tag;
        return !empty && !dup;
    };
    uint32_t local = 0;
    for (uint32_t i = s0; i < s1; ++i) local += keep(i);
    psum[tid] = local;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const uint32_t v = tid >= off ? psum[tid - off] : 0;
        __syncthreads();
        psum[tid] += v;
        __syncthreads();
    }
    uint32_t pos = psum[tid] - local;
    for (uint32_t i = s0; i < s1 && pos < N; ++i)
        if (keep(i)) top[pos++] = work[i];
    __threadfence_block();
    __syncthreads();
    if (tid == 0) {
        const uint32_t n_top = min(N, psum[1023]);
        st->n_top = n_top;
        if (n_top >= N && N > 0) {
            st->tau = top[N - 1].mi;
            st->tau_combo = top[N - 1].combo;
        }
        st->cand_count = 0;
        if (st->overflow) {
            redo[st->redo_count++] = chunk_id;
            st->overflow = 0;
        }
    }
}

__global__ void __launch_bounds__(1024)
k_merge(TopState *st, Cand *top, const Cand *cand, uint32_t N, uint32_t cap, uint32_t chunk_id,
        uint32_t *redo)
{
    __shared__ Cand buf[kMergeSlots];
    __shared__ uint32_t psum[1024];
    merge_topn(st, top, cand, N, cap, buf, psum);
    if (threadIdx.x == 0 && st->overflow) {
        redo[st->redo_count++] = chunk_id;
        st->overflow = 0;
    }
}

// ---------------------------------------------------------------------------
// Batched scans.
// ---------------------------------------------------------------------------
// Persistent scan over the segments of one launch: every wavefront takes the
// next unit from a counter until the units run out (every wave reaches the
// exit, so the grid drains).  LAUNCH 0 runs the order-2/3 kinds in one
// kernel, LAUNCH 1 the direct order-4 kind and LAUNCH 2 the marginal
// order-4/5 kinds, whose register budgets would otherwise set the occupancy
// of the others.
#if defined(__HIP_DEVICE_COMPILE__)
typedef const __attribute__((address_space(4))) BatchJob ConstJob;
#else
typedef const BatchJob ConstJob;  // host pass: kernel bodies are only parsed
#endif

template <int LAUNCH>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(LAUNCH == 0 ? 3 : LAUNCH == 2 ? 2 : 1)))
k_batch_scan(const BatchJob *__restrict__ jobs, const BatchSeg *__restrict__ segs, BatchHeader *hdr)
{
    const int lane = threadIdx.x & 63;
    const uint32_t nseg = hdr->nseg[LAUNCH], units = hdr->units[LAUNCH];
    for (;;) {
        uint32_t u = 0;
        if (lane == 0) u = atomicAdd(&hdr->head[LAUNCH], 1u);
        u = __builtin_amdgcn_readfirstlane(u);
        if (u >= units) return;
        // segment holding unit u: the last one whose start is <= u
        uint32_t lo = 0, hi = nseg - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (segs[mid].start <= u) lo = mid;
            else hi = mid - 1;
        }
        const int64_t j0 = segs[lo].j0, j1 = segs[lo].j1;
        // The job table is read-only while the launch runs: reading it through
        // the constant address space lets the compiler treat every field as
        // invariant (scalar loads it may repeat instead of keeping registers).
        const ConstJob &J = ((const ConstJob *)jobs)[segs[lo].job];
        const DevData D = J.D;
        const ScanParams P = J.P;
        TopState *st = J.st;
        const uint32_t chunks = (uint32_t)J.chunks, local = u - (uint32_t)segs[lo].start;
        const int64_t g = (int64_t)(local / chunks);
        const int c = (int)(local % chunks);
        const int t_hi = (D.n_pad >> 6) - 1 - c * J.tiles, t_lo = t_hi - J.tiles + 1;
        if constexpr (LAUNCH == 0) {
            switch (J.kind) {
            case kKindScan2m:
                scan2m_unit<kMarginalPW, false, false>(D, P, st, nullptr, j0 + g * kMarginalPW, j1, t_hi, t_lo,
                                                       nullptr);
                break;
            case kKindScan3m:
                scan3m_unit<kMarginalPW, false>(D, P, st, j0 + g * kMarginalPW, j1, t_hi, t_lo, nullptr);
                break;
            case kKindDirect2: direct_unit<2, false>(D, P, st, j0 + g, j1, t_hi, t_lo, nullptr); break;
            case kKindDirect3: direct_unit<3, false>(D, P, st, j0 + g, j1, t_hi, t_lo, nullptr); break;
            default: break;
            }
        } else if constexpr (LAUNCH == 1) {
            direct_unit<4, false>(D, P, st, j0 + g, j1, t_hi, t_lo, nullptr);
        } else {
            if (J.kind == kKindKm4) km_unit<4, false>(D, P, st, j0 + g, j1, t_hi, t_lo, nullptr);
            else if (J.kind == kKindKm5) km_unit<5, false>(D, P, st, j0 + g, j1, t_hi, t_lo, nullptr);
            else km_unit<5, false, true>(D, P, st, j0 + g, j1, t_hi, t_lo, nullptr);
        }
    }
}

// Matrix-core kinds of a batched step: persistent workgroups with one work
// queue per XCD (workgroup b runs on XCD b % 8 -- the dispatch order -- and
// takes units from queue b % 8; every queue has 1/8 of the workgroups, so all
// are drained wherever they run).  Queue x's unit v lies in the segment whose
// start is the last <= v; with l = v - start, order 3: first SNP j0 + l % na,
// tile 8 (l / na) + x (the XCD-ordered work of k_mfma3w / k_mfma3); order 2:
// super-block j0 + 8 (l / 32) + x, 64 x 128 tile l % 32 (k_mfma2w), or 8 (l /
// 64) + x, 64 x 64 tile l % 64 (k_mfma2).  The job's `wide` picks the tiles.
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_batch_mfma(const BatchJob *__restrict__ jobs, const BatchSeg *__restrict__ segs, BatchHeader *hdr)
{
    constexpr size_t kWide = (size_t)kM2wRing * kM2wStage > (size_t)kM3Ring * kM3wStage
                                 ? (size_t)kM2wRing * kM2wStage
                                 : (size_t)kM3Ring * kM3wStage;
    constexpr size_t kNarrow = sizeof(Mfma2Lds<true>) > (size_t)kM3Ring * kM3Stage ? sizeof(Mfma2Lds<true>)
                                                                                 : (size_t)kM3Ring * kM3Stage;
    constexpr size_t kBytes = kWide > kNarrow ? kWide : kNarrow;
    __shared__ __attribute__((aligned(1024))) int8_t lds[kBytes];   // the order-2 or the order-3 tiles
    __shared__ uint32_t unit_sh;
    const int x = blockIdx.x & 7;
    const uint32_t nseg = hdr->mseg, units = hdr->munits;
    for (;;) {
        if (threadIdx.x == 0) unit_sh = atomicAdd(&hdr->mhead[x], 1u);
        __syncthreads();
        const uint32_t v = __builtin_amdgcn_readfirstlane(unit_sh);
        __syncthreads();   // unit_sh is rewritten only after every wave has read it
        if (v >= units) return;
        uint32_t lo = 0, hi = nseg - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (segs[mid].start <= v) lo = mid;
            else hi = mid - 1;
        }
        const ConstJob &J = ((const ConstJob *)jobs)[segs[lo].job];
        const DevData D = J.D;
        const ScanParams P = J.P;
        const Fp4Image F = J.F;
        TopState *st = J.st;
        const uint32_t l = v - (uint32_t)segs[lo].start;
        const int64_t j0 = segs[lo].j0, j1 = segs[lo].j1;
        if (J.kind == kKindMfma3) {
            const uint32_t na = (uint32_t)(j1 - j0);
            if (J.wide)
                mfma3w_tile<kM3Ring>(D, P, st, F, (int)(j0 + l % na), (int64_t)(l / na) * 8 + x, J.T,
                                     (int8_t(*)[kM3wStage])lds);
            else
                mfma3_tile<kM3Ring>(D, P, st, F, (int)(j0 + l % na), (int64_t)(l / na) * 8 + x, J.T,
                                    (int8_t(*)[kM3Stage])lds);
        } else if (J.wide) {
            const int64_t su = j0 + (int64_t)(l / 32) * 8 + x;
            if (su < j1) mfma2w_tile<kBatchM2wKeep>(D, P, st, F, su, (int)(l % 32), J.T, (int8_t(*)[kM2wStage])lds);
        } else {
            const int64_t su = j0 + (int64_t)(l / 64) * 8 + x;
            if (su < j1) mfma2_tile<true>(D, P, st, F, su, (int)(l % 64), J.T, *(Mfma2Lds<true> *)lds);
        }
    }
}

// One workgroup per job of the launch: merge its candidates into its top-N
// and record whether its candidate buffer overflowed in this launch.
__global__ void __launch_bounds__(1024)
k_batch_merge(const BatchJob *__restrict__ jobs, const int32_t *__restrict__ list, const BatchHeader *hdr)
{
    __shared__ Cand buf[kMergeSlots];
    __shared__ uint32_t psum[1024];
    if (blockIdx.x >= hdr->nmerge) return;
    const BatchJob &J = jobs[list[blockIdx.x]];
    merge_topn(J.st, J.top, J.P.cand, J.outputs, J.P.cand_cap, buf, psum);
    if (threadIdx.x == 0) {
        J.st->step_overflow = J.st->overflow;
        J.st->overflow = 0;
    }
}

__global__ void k_batch_reset(BatchHeader *hdr)
{
    if (threadIdx.x == 0) {
        hdr->head[0] = 0;
        hdr->head[1] = 0;
        hdr->head[2] = 0;
    }
    if (threadIdx.x < 8) hdr->mhead[threadIdx.x] = 0;
}

// ---------------------------------------------------------------------------
// Dataset ingest: genotype matrix -> reference-layout words.
// ---------------------------------------------------------------------------
// A workgroup packs whole SNPs (grid-stride).  With `Stage`, the SNP's
// genotype row is first copied into LDS with aligned 16-byte loads and its
// 3(cw+tw) output words are assembled in LDS and written out with 16-byte
// stores; without it (rows too long for LDS) both go straight to memory.
// One wavefront per word slot (class c, word w), all three rows: lane i takes
// the genotype of the class's sample 64w+i -- sample start_c + 64w + i of the
// file when each class is one contiguous run of samples, else the one the
// per-class sample list `idx` names (cases then controls, each in file
// order) -- a ballot per genotype value puts lane i at bit i, and the bit
// reversal puts sample 64w at the most significant bit (RAWFile.hpp:233-260:
// buff = (buff<<1) + bit).  A last partial word of r samples is right-aligned
// (RAWFile.hpp:263-283): the reversed ballot shifted right by 64 - r.
// Padding words, and the slots flagged in `zero` (the readers' second flush
// of an emptied buffer), are 0.
struct PackArgs {
    const int8_t *geno;      // snps x samples, allocated to a multiple of 16 bytes
    const int32_t *idx;      // per-class sample lists, or null: contiguous classes
    const uint8_t *zero;     // per word slot (cw + tw), or null
    uint64_t *words;         // snps x 3(cw + tw)
    uint64_t samples;
    uint32_t n_class[2], start[2];
    uint32_t cw, tw, snps;
};

template <bool Stage>
__global__ __launch_bounds__(256) void k_pack_words(PackArgs A)
{
    extern __shared__ __align__(16) uint8_t pack_lds[];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t per = A.cw + A.tw;
    uint64_t *out_lds = (uint64_t *)pack_lds;   // 3 * per words, then the row
    for (uint32_t s = blockIdx.x; s < A.snps; s += gridDim.x) {
        const uint64_t r0 = (uint64_t)s * A.samples;
        const int8_t *row;
        uint64_t *out;
        if constexpr (Stage) {
            const uint64_t a0 = r0 & ~(uint64_t)15, a1 = (r0 + A.samples + 15) & ~(uint64_t)15;
            uint4 *buf = (uint4 *)(pack_lds + (size_t)per * 24);
            const uint4 *src = (const uint4 *)(A.geno + a0);
            __syncthreads();   // the previous SNP's words are out
            for (uint32_t i = threadIdx.x; i < (uint32_t)((a1 - a0) >> 4); i += blockDim.x) buf[i] = src[i];
            __syncthreads();
            row = (const int8_t *)buf + (r0 - a0);
            out = out_lds;
        } else {
            row = A.geno + r0;
            out = A.words + (uint64_t)s * 3 * per;
        }
        for (uint32_t slot = wave; slot < per; slot += blockDim.x >> 6) {
            const int c = slot >= A.cw;
            const uint32_t w = c ? slot - A.cw : slot, W = c ? A.tw : A.cw;
            const int64_t have = (int64_t)A.n_class[c] - 64 * (int64_t)w;   // samples in this word
            int g = -1;
            if (lane < have) {
                const uint32_t pos = 64 * w + lane;
                const uint64_t j = A.idx ? (uint64_t)A.idx[(c ? A.n_class[0] : 0) + pos] : (uint64_t)A.start[c] + pos;
                g = row[j];
            }
            uint64_t v[3];
#pragma unroll
            for (int z = 0; z < 3; ++z) {
                const uint64_t m = __builtin_bitreverse64(__ballot(g == z));
                v[z] = have >= 64 ? m : have > 0 ? m >> (64 - have) : 0;
            }
            if (A.zero && A.zero[slot]) v[0] = v[1] = v[2] = 0;
            if (lane < 3) out[(c ? 3 * A.cw : 0) + lane * W + w] = lane == 0 ? v[0] : lane == 1 ? v[1] : v[2];
        }
        if constexpr (Stage) {
            __syncthreads();
            const uint4 *o = (const uint4 *)out_lds;
            uint4 *dst = (uint4 *)(A.words + (uint64_t)s * 3 * per);   // 3 * per * 8 bytes: a multiple of 16
            for (uint32_t i = threadIdx.x; i < per * 3 / 2; i += blockDim.x) dst[i] = o[i];
        }
    }
}

// ---------------------------------------------------------------------------
// Dataset image: transpose + per-row counts + coverage check.
// ---------------------------------------------------------------------------
__global__ void k_transpose(const uint64_t *__restrict__ src, uint4 *__restrict__ dst, int n, int n_pad,
                            int cls, uint64_t cw, uint64_t tw, int quads, uint64_t class_off)
{
    // one thread per (q, z, s)
    const uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t total = (uint64_t)quads * 3 * n_pad;
    if (id >= total) return;
    const int s = (int)(id % n_pad);
    const uint64_t qz = id / n_pad;
    const int z = (int)(qz % 3), q = (int)(qz / 3);
    uint4 v = make_uint4(0, 0, 0, 0);
    if (s < n) {
        const uint64_t words = cls == 0 ? cw : tw;
        const uint64_t *row = src + (uint64_t)s * 3 * (cw + tw) + (cls == 0 ? 0 : 3 * cw) + (uint64_t)z * words;
        const uint64_t w0 = 2 * (uint64_t)q, w1 = w0 + 1;
        const uint64_t lo = w0 < words ? row[w0] : 0, hi = w1 < words ? row[w1] : 0;
        v = make_uint4((uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32));
    }
    dst[class_off + id] = v;
}

__global__ void k_singles(const uint4 *__restrict__ bits, DevData D, uint32_t *singles, uint32_t *flags)
{
    // one thread per (SNP, class): popcount of each row, disjointness of the
    // rows and equality of their union with the union of SNP 0.
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= D.n_pad * 2) return;
    const int s = id >> 1, cls = id & 1;
    const uint4 *base = bits + D.class_off[cls];
    const int np = D.n_pad;
    uint32_t cnt[3] = {0, 0, 0}, bad = 0, ucnt = 0;
    for (int q = 0; q < D.quads[cls]; ++q) {
        const uint4 *row = base + (size_t)q * 3 * np;
        const uint4 r0 = row[s], r1 = row[np + s], r2 = row[2 * np + s];
        const uint4 u0 = row[0], u1 = row[np], u2 = row[2 * np];
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            const uint32_t a = comp(r0, w), b = comp(r1, w), c = comp(r2, w);
            const uint32_t ref = comp(u0, w) | comp(u1, w) | comp(u2, w);
            cnt[0] += __builtin_popcount(a);
            cnt[1] += __builtin_popcount(b);
            cnt[2] += __builtin_popcount(c);
            bad |= (a & b) | (a & c) | (b & c) | ((a | b | c) ^ ref);
            ucnt += __builtin_popcount(ref);
        }
    }
    singles[(s * 2 + cls) * 3 + 0] = cnt[0];
    singles[(s * 2 + cls) * 3 + 1] = cnt[1];
    singles[(s * 2 + cls) * 3 + 2] = cnt[2];
    if (s < D.n && bad) atomicOr(&flags[0], 1u);
    if (s == 0) flags[1 + cls] = ucnt;
}

// ---------------------------------------------------------------------------
// Batch operators.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t load_word(const uint4 *bits, const DevData &D, int cls, int z, int s, uint64_t w)
{
    const uint4 v = bits[D.class_off[cls] + ((w >> 1) * 3 + z) * (uint64_t)D.n_pad + s];
    return (w & 1) ? ((uint64_t)v.z | ((uint64_t)v.w << 32)) : ((uint64_t)v.x | ((uint64_t)v.y << 32));
}

__global__ void k_combine_batch(const uint4 *__restrict__ bits, DevData D, const int32_t *combos, int k,
                                uint64_t n, uint64_t cw, uint64_t tw, uint64_t *out)
{
    uint64_t R = 1;
    for (int i = 0; i < k; ++i) R *= 3;
    const uint64_t per = R * (cw + tw);
    const uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n * per) return;
    const uint64_t comb = id / per, rem = id % per;
    const int cls = rem < R * cw ? 0 : 1;
    const uint64_t r2 = cls == 0 ? rem : rem - R * cw, W = cls == 0 ? cw : tw;
    const uint64_t r = r2 / W, w = r2 % W;
    uint64_t v = ~0ull, rr = r;
    for (int i = k - 1; i >= 0; --i) {
        const int z = (int)(rr % 3);
        rr /= 3;
        v &= load_word(bits, D, cls, z, combos[comb * k + i], w);
    }
    out[id] = v;
}

__global__ void k_contingency_batch(const uint4 *__restrict__ bits, DevData D, const int32_t *combos, int k,
                                    uint64_t ct_size, uint32_t *out)
{
    // one wavefront per combination; lanes split the words, shuffles reduce
    const uint64_t comb = blockIdx.x;
    const int lane = threadIdx.x;
    int c[16];
    for (int i = 0; i < k; ++i) c[i] = combos[comb * k + i];
    uint64_t cells = 1;
    for (int i = 0; i < k; ++i) cells *= 3;
    for (int cls = 0; cls < 2; ++cls) {
        const uint64_t words = 4ull * D.quads[cls];
        for (uint64_t cell = 0; cell < ct_size; ++cell) {
            uint32_t acc = 0;
            if (cell < cells) {
                for (uint64_t w = lane; w < words; w += 64) {
                    uint32_t v = ~0u;
                    uint64_t rr = cell;
                    for (int i = k - 1; i >= 0; --i) {
                        const int z = (int)(rr % 3);
                        rr /= 3;
                        const uint4 q = bits[D.class_off[cls] + ((w >> 2) * 3 + z) * (uint64_t)D.n_pad + c[i]];
                        v &= comp(q, (int)(w & 3));
                    }
                    acc += __builtin_popcount(v);
                }
            }
            for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
            if (lane == 0) out[comb * 2 * ct_size + cls * ct_size + cell] = acc;
        }
    }
}

__global__ void k_mi_batch(const uint32_t *tables, uint64_t n, uint64_t ct_size, double n64, double hy64,
                           float *out)
{
#pragma clang fp contract(off)
    const uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const uint32_t *a0 = tables + id * 2 * ct_size, *a1 = a0 + ct_size;
    double hx = 0.0, hall = 0.0;
    for (uint64_t x = 0; x < ct_size; ++x) {
        const double pc = (double)a0[x] / n64, pt = (double)a1[x] / n64, pa = pc + pt;
        if (pc > 0.0) hall -= pc * log(pc);
        if (pt > 0.0) hall -= pt * log(pt);
        if (pa > 0.0) hx -= pa * log(pa);
    }
    out[id] = (float)(hx + hy64 - hall);
}

// ---------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------
// A dispatch holds at most 2^32 - 1 work-items per grid dimension: launches
// over prefix ranges are split into ranges whose grids stay below that.
constexpr int64_t kMaxGridBlocks = (int64_t)(0xffffffffu / (uint32_t)kBlock);

// f(Q) for consecutive sub-ranges [Q.j0, Q.j1) of [P.j0, P.j1) of at most
// kMaxGridBlocks x per_block prefixes each.
template <class F>
static hipError_t split_prefixes(const ScanParams &P, int64_t per_block, F &&f)
{
    const int64_t span = kMaxGridBlocks * per_block;
    for (int64_t j = P.j0; j < P.j1; j += span) {
        ScanParams Q = P;
        Q.j0 = j;
        Q.j1 = std::min<int64_t>(P.j1, j + span);
        const hipError_t e = f(Q);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <bool HIST>
static hipError_t scan_direct(int order, const DevData &D, const ScanParams &P0, TopState *st, hipStream_t s)
{
    if (P0.j1 - P0.j0 > (order >= 7 && order <= 9 ? 1 : 4) * kMaxGridBlocks && (order < 10 || order > 12))
        return split_prefixes(P0, order >= 7 ? 1 : 4,
                              [&](const ScanParams &Q) { return scan_direct<HIST>(order, D, Q, st, s); });
    const ScanParams &P = P0;
    const int64_t waves = P.j1 - P.j0;
    if (waves <= 0) return hipSuccess;
    const dim3 grid((unsigned)((waves + 3) / 4)), block(kBlock);
    switch (order) {
    case 2: hipLaunchKernelGGL((k_scan_direct<2, HIST>), grid, block, 0, s, D, P, st); break;
    case 3: hipLaunchKernelGGL((k_scan_direct<3, HIST>), grid, block, 0, s, D, P, st); break;
    case 4: hipLaunchKernelGGL((k_scan_direct<4, HIST>), grid, block, 0, s, D, P, st); break;
    // one prefix per workgroup
    case 7: hipLaunchKernelGGL((k_scan_any<7, HIST>), dim3((unsigned)waves), block, 0, s, D, P, st); break;
    case 8: hipLaunchKernelGGL((k_scan_any<8, HIST>), dim3((unsigned)waves), block, 0, s, D, P, st); break;
    case 9: hipLaunchKernelGGL((k_scan_any<9, HIST>), dim3((unsigned)waves), block, 0, s, D, P, st); break;
    // persistent, one count slot of P.cells per workgroup
    case 10: case 11: case 12: {
        if (!P.cells) return hipErrorInvalidValue;
        const dim3 g((unsigned)std::min<int64_t>(waves, big_slots(order)));
        if (order == 10) hipLaunchKernelGGL((k_scan_big<10, HIST>), g, block, 0, s, D, P, st, P.cells);
        else if (order == 11) hipLaunchKernelGGL((k_scan_big<11, HIST>), g, block, 0, s, D, P, st, P.cells);
        else hipLaunchKernelGGL((k_scan_big<12, HIST>), g, block, 0, s, D, P, st, P.cells);
        break;
    }
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_scan_km(int order, const DevData &D, const ScanParams &P, TopState *st, hipStream_t s,
                          bool all_cells)
{
    const int64_t prefixes = P.j1 - P.j0;
    if (prefixes <= 0) return hipSuccess;
    const int64_t per_block = kBlock / kWave;
    if (prefixes > per_block * kMaxGridBlocks)
        return split_prefixes(P, per_block,
                              [&](const ScanParams &Q) { return launch_scan_km(order, D, Q, st, s, all_cells); });
    const dim3 grid((unsigned)((prefixes + per_block - 1) / per_block)), block(kBlock);
    const bool hist = P.hist != nullptr;
#define FIU_KM(K, H, A) hipLaunchKernelGGL((k_scan_km<K, H, A>), grid, block, 0, s, D, P, st)
#define FIU_KM_ORDER(K)                                                     \
    if (all_cells) {                                                       \
        if (hist) FIU_KM(K, true, true);                                   \
        else FIU_KM(K, false, true);                                       \
    } else {                                                               \
        if (hist) FIU_KM(K, true, false);                                  \
        else FIU_KM(K, false, false);                                      \
    }
    switch (order) {
    case 4: FIU_KM_ORDER(4) break;
    case 5: FIU_KM_ORDER(5) break;
    case 6: FIU_KM_ORDER(6) break;
    default: return hipErrorInvalidValue;
    }
#undef FIU_KM_ORDER
#undef FIU_KM
    return hipGetLastError();
}

hipError_t launch_scan_direct(int order, const DevData &D, const ScanParams &P, TopState *st, hipStream_t s)
{
    // orders 5 and 6: every cell, the prefix rows split over lanes
    if (order == 5 || order == 6) return launch_scan_km(order, D, P, st, s, true);
    return P.hist ? scan_direct<true>(order, D, P, st, s) : scan_direct<false>(order, D, P, st, s);
}

template <int PW, bool HIST>
static hipError_t scan_marginal(int order, const DevData &D, const ScanParams &P, TopState *st, hipStream_t s)
{
    const int64_t prefixes = P.j1 - P.j0;
    if (prefixes <= 0) return hipSuccess;
    const int64_t per_block = (kBlock / kWave) * PW;
    if (prefixes > per_block * kMaxGridBlocks)
        return split_prefixes(P, per_block,
                              [&](const ScanParams &Q) { return scan_marginal<PW, HIST>(order, D, Q, st, s); });
    const dim3 grid((unsigned)((prefixes + per_block - 1) / per_block)), block(kBlock);
    switch (order) {
    case 2: hipLaunchKernelGGL((k_scan2m<PW, false, HIST>), grid, block, 0, s, D, P, st, (uint32_t *)nullptr); break;
    case 3:
        if constexpr (PW <= kMarginalPW) {   // 8 prefixes per wave do not fit order 3's registers
            hipLaunchKernelGGL((k_scan3m<PW, HIST>), grid, block, 0, s, D, P, st);
            break;
        }
        return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_scan_marginal(int order, const DevData &D, const ScanParams &P, TopState *st, hipStream_t s,
                                int pw)
{
    if (P.hist) return scan_marginal<kMarginalPW, true>(order, D, P, st, s);
    switch (pw) {
    case 2: return scan_marginal<2, false>(order, D, P, st, s);
    case kMarginalPW: return scan_marginal<kMarginalPW, false>(order, D, P, st, s);
    case 8: return scan_marginal<8, false>(order, D, P, st, s);
    default: return hipErrorInvalidValue;
    }
}

int batch_blocks(int launch)
{
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    const hipError_t e =
        launch == 0   ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_scan<0>, kBlock, 0)
        : launch == 1 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_scan<1>, kBlock, 0)
        : launch == 3 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_scan<2>, kBlock, 0)
                      : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_mfma, kBlock, 0);
    if (e != hipSuccess) return 0;
    const int blocks = std::max(1, per_cu) * cus;
    return launch == 2 ? std::max(8, blocks / 8 * 8) : blocks;   // matrix-core queues: a multiple of 8
}

hipError_t launch_batch(const BatchJob *jobs, const BatchSeg *segs0, const BatchSeg *segs1, const BatchSeg *segs2,
                        const BatchSeg *segs3, const int32_t *merge_list, BatchHeader *hdr, int blocks0, int blocks1,
                        int blocks2, int blocks3, int merge_blocks, hipStream_t s)
{
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, hdr);
    hipLaunchKernelGGL((k_batch_scan<0>), dim3(blocks0), dim3(kBlock), 0, s, jobs, segs0, hdr);
    if (blocks1 > 0) hipLaunchKernelGGL((k_batch_scan<1>), dim3(blocks1), dim3(kBlock), 0, s, jobs, segs1, hdr);
    if (blocks3 > 0) hipLaunchKernelGGL((k_batch_scan<2>), dim3(blocks3), dim3(kBlock), 0, s, jobs, segs3, hdr);
    if (blocks2 > 0) hipLaunchKernelGGL(k_batch_mfma, dim3(blocks2), dim3(kBlock), 0, s, jobs, segs2, hdr);
    hipLaunchKernelGGL(k_batch_merge, dim3(merge_blocks), dim3(1024), 0, s, jobs, merge_list, hdr);
    return hipGetLastError();
}

hipError_t launch_tau_from_hist(const uint32_t *hist, uint32_t N, float delta, TopState *st, hipStream_t s)
{
    hipLaunchKernelGGL(k_tau_from_hist, dim3(1), dim3(kTauThreads), 0, s, hist, N, delta, st);
    return hipGetLastError();
}

int64_t mfma2_superblocks(int T)
{
    const int64_t S = (T + kMfmaSuper - 1) / kMfmaSuper;
    return S * (S + 1) / 2;
}

hipError_t launch_mfma2(const DevData &D, const ScanParams &P, TopState *st, int64_t s0, int64_t s1, int T,
                        const uint4 *fp4, int tile, hipStream_t s)
{
    if (s1 <= s0) return hipSuccess;
    if (!fp4) tile = 0;
    const int per = tile == 2 ? 16 : tile == 1 ? 32 : 64;   // blocks per super-block
    // a grid below 2^32 work-items (blocks of 2 kBlock for the 128 x 128
    // tile): sub-launches of at most this many super-blocks
    const int64_t max_super = kMaxGridBlocks / (8 * per * (tile == 2 ? 2 : 1)) * 8;
    if (s1 - s0 > max_super) {
        for (int64_t a = s0; a < s1; a += max_super) {
            const hipError_t e = launch_mfma2(D, P, st, a, std::min(s1, a + max_super), T, fp4, tile, s);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    // `per` blocks per super-block, 8 super-blocks (one per XCD) per round
    const int64_t rounds = (s1 - s0 + 7) / 8;
    const dim3 grid((unsigned)(rounds * 8 * per)), block(tile == 2 ? 2 * kBlock : kBlock);
    Fp4Image F;
    F.base = fp4;
    F.class_off[0] = 0;
    F.class_off[1] = fp4_image_units(D, 0);
    if (tile == 2) hipLaunchKernelGGL(k_mfma2x, grid, block, 0, s, D, P, st, F, s0, s1, T);
    else if (tile == 1) hipLaunchKernelGGL(k_mfma2w, grid, block, 0, s, D, P, st, F, s0, s1, T);
    else if (fp4) hipLaunchKernelGGL((k_mfma2<true>), grid, block, 0, s, D, P, st, F, s0, s1, T);
    else hipLaunchKernelGGL((k_mfma2<false>), grid, block, 0, s, D, P, st, F, s0, s1, T);
    return hipGetLastError();
}

uint64_t fp4_image_units(const DevData &D, int cls) { return 8ull * D.quads[cls] * D.n_pad; }

hipError_t launch_fp4_image(const DevData &D, uint4 *out, hipStream_t s)
{
    const uint64_t total = 2ull * (D.quads[0] + D.quads[1]) * D.n_pad;
    if (!total) return hipSuccess;
    hipLaunchKernelGGL(k_fp4_image, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, D.bits, D, out, 0ull,
                       fp4_image_units(D, 0));
    return hipGetLastError();
}

hipError_t launch_mfma3(const DevData &D, const ScanParams &P, TopState *st, const uint4 *fp4, int a0, int a1, int T,
                        int ring, hipStream_t s)
{
    if (a1 <= a0) return hipSuccess;
    const int m = T - ((a0 + 1) >> 6);   // tiles from the first SNP's tile on (the most for a0)
    if (m <= 0) return hipSuccess;
    Fp4Image F;
    F.base = fp4;
    F.class_off[0] = 0;
    F.class_off[1] = fp4_image_units(D, 0);
    // work items per first SNP: 64 x 32 half tiles, or whole 64 x 64 tiles
    // (rings 5 and 6: the wide tile with 3 and 4 stages)
    const bool wide = ring >= 5;
    const int64_t items = wide ? (int64_t)m * (m + 1) / 2 : (int64_t)m * (m + 1);
    const int64_t tiles8 = (items + 7) / 8, na = a1 - a0;
    if (tiles8 * na * 8 > kMaxGridBlocks) {
        // a grid below 2^32 work-items: fewer first SNPs per launch
        const int64_t step = std::max<int64_t>(1, kMaxGridBlocks / (tiles8 * 8));
        for (int64_t a = a0; a < a1; a += step) {
            const hipError_t e = launch_mfma3(D, P, st, fp4, (int)a, (int)std::min<int64_t>(a1, a + step), T, ring, s);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    const dim3 grid((unsigned)(tiles8 * na * 8)), block(kBlock);
    if (ring == 3) hipLaunchKernelGGL((k_mfma3<3>), grid, block, 0, s, D, P, st, F, a0, (int)na, T);
    else if (ring == kM3Ring) hipLaunchKernelGGL((k_mfma3<kM3Ring>), grid, block, 0, s, D, P, st, F, a0, (int)na, T);
    else if (ring == 5) hipLaunchKernelGGL((k_mfma3w<3>), grid, block, 0, s, D, P, st, F, a0, (int)na, T);
    else if (ring == 6) hipLaunchKernelGGL((k_mfma3w<4>), grid, block, 0, s, D, P, st, F, a0, (int)na, T);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_pair_table(const DevData &D, const ScanParams &P, uint32_t *pair_out, hipStream_t s)
{
    const int64_t prefixes = P.j1 - P.j0;
    if (prefixes <= 0) return hipSuccess;
    const int64_t per_block = (kBlock / kWave) * kMarginalPW;
    const int64_t blocks = (prefixes + per_block - 1) / per_block;
    // at least about 2,048 wavefronts: a small dataset's prefixes alone leave
    // most CUs idle, so the tiles of last SNPs are split as well
    const int T = D.n_pad >> 6;
    const int chunks = (int)std::min<int64_t>(T, std::max<int64_t>(1, 512 / std::max<int64_t>(blocks, 1)));
    const dim3 grid((unsigned)blocks, (unsigned)chunks), block(kBlock);
    hipLaunchKernelGGL((k_scan2m<kMarginalPW, true, false>), grid, block, 0, s, D, P, (TopState *)nullptr, pair_out);
    return hipGetLastError();
}

hipError_t launch_merge(TopState *st, Cand *top, const Cand *cand, uint32_t N, uint32_t cap, uint32_t chunk_id,
                        uint32_t *redo, hipStream_t s)
{
    hipLaunchKernelGGL(k_merge, dim3(1), dim3(1024), 0, s, st, top, cand, N, cap, chunk_id, redo);
    return hipGetLastError();
}

hipError_t launch_pack_words(const int8_t *geno, uint64_t snps, uint64_t samples, const int32_t *idx,
                             const uint32_t n_class[2], const uint32_t start[2], uint64_t cw, uint64_t tw,
                             const uint8_t *zero, uint64_t *words, hipStream_t s)
{
    if (snps == 0 || cw + tw == 0) return hipSuccess;
    PackArgs A;
    A.geno = geno;
    A.idx = idx;
    A.zero = zero;
    A.words = words;
    A.samples = samples;
    for (int c = 0; c < 2; ++c) A.n_class[c] = n_class[c], A.start[c] = start[c];
    A.cw = (uint32_t)cw;
    A.tw = (uint32_t)tw;
    A.snps = (uint32_t)snps;
    const unsigned blocks = (unsigned)std::min<uint64_t>(snps, 8192);
    const size_t lds = pack_lds_bytes(samples, cw, tw);
    if (lds <= kPackLdsMax)
        hipLaunchKernelGGL(k_pack_words<true>, dim3(blocks), dim3(256), lds, s, A);
    else
        hipLaunchKernelGGL(k_pack_words<false>, dim3(blocks), dim3(256), 0, s, A);
    return hipGetLastError();
}

hipError_t launch_transpose(const uint64_t *src, uint4 *dst, int n, int n_pad, int cls, uint64_t cw, uint64_t tw,
                            int quads, uint64_t class_off, hipStream_t s)
{
    const uint64_t total = (uint64_t)quads * 3 * n_pad;
    if (!total) return hipSuccess;
    hipLaunchKernelGGL(k_transpose, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, src, dst, n, n_pad, cls,
                       cw, tw, quads, class_off);
    return hipGetLastError();
}

hipError_t launch_singles(const uint4 *bits, const DevData &D, uint32_t *singles, uint32_t *flags, hipStream_t s)
{
    const int total = D.n_pad * 2;
    hipLaunchKernelGGL(k_singles, dim3((total + 255) / 256), dim3(256), 0, s, bits, D, singles, flags);
    return hipGetLastError();
}

hipError_t launch_combine_batch(const DevData &D, const int32_t *combos, int k, uint64_t n, uint64_t cw, uint64_t tw,
                                uint64_t *out, hipStream_t s)
{
    uint64_t R = 1;
    for (int i = 0; i < k; ++i) R *= 3;
    const uint64_t per = R * (cw + tw), total = n * per;
    if (!total) return hipSuccess;
    // grids below 2^32 work-items: whole combinations per launch
    const uint64_t max_n = std::max<uint64_t>(1, 0xffffff00ull / per);
    if (per > 0xffffff00ull) return hipErrorInvalidValue;
    for (uint64_t c = 0; c < n; c += max_n) {
        const uint64_t m = std::min(n - c, max_n);
        hipLaunchKernelGGL(k_combine_batch, dim3((unsigned)((m * per + 255) / 256)), dim3(256), 0, s, D.bits, D,
                           combos + c * k, k, m, cw, tw, out + c * per);
    }
    return hipGetLastError();
}

hipError_t launch_contingency_batch(const DevData &D, const int32_t *combos, int k, uint64_t n, uint64_t ct_size,
                                    uint32_t *out, hipStream_t s)
{
    // grids below 2^32 work-items (one 64-thread workgroup per combination)
    constexpr uint64_t kMaxN = 0xffffffffull / 64;
    for (uint64_t c = 0; c < n; c += kMaxN) {
        const uint64_t m = std::min(n - c, kMaxN);
        hipLaunchKernelGGL(k_contingency_batch, dim3((unsigned)m), dim3(64), 0, s, D.bits, D, combos + c * k, k,
                           ct_size, out + c * 2 * ct_size);
    }
    return hipGetLastError();
}

hipError_t launch_mi_batch(const uint32_t *tables, uint64_t n, uint64_t ct_size, double n64, double hy64, float *out,
                           hipStream_t s)
{
    // grids below 2^32 work-items
    constexpr uint64_t kMaxN = 0xffffff00ull;
    for (uint64_t c = 0; c < n; c += kMaxN) {
        const uint64_t m = std::min(n - c, kMaxN);
        hipLaunchKernelGGL(k_mi_batch, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, tables + c * 2 * ct_size, m,
                           ct_size, n64, hy64, out + c);
    }
    return hipGetLastError();
}

hipError_t launch_merge_big(TopState *st, Cand *top, uint32_t n_top, const Cand *cand, uint32_t count, uint32_t N,
                            Cand *work, uint32_t L, uint32_t chunk_id, uint32_t *redo, hipStream_t s)
{
    if (L < 2 || (L & (L - 1)) || L < n_top + count) return hipErrorInvalidValue;
    const unsigned blocks = (L + 255) / 256;
    hipLaunchKernelGGL(k_big_gather, dim3(blocks), dim3(256), 0, s, top, n_top, cand, count, work, L);
    for (uint32_t k = 2; k <= L; k <<= 1)
        for (uint32_t jj = k >> 1; jj > 0; jj >>= 1)
            hipLaunchKernelGGL(k_big_bitonic, dim3(blocks), dim3(256), 0, s, work, L, k, jj);
    hipLaunchKernelGGL(k_big_finish, dim3(1), dim3(1024), 0, s, work, L, top, N, st, chunk_id, redo);
    return hipGetLastError();
}

hipError_t launch_binom_table(uint64_t *binom, uint64_t n, int m, hipStream_t s)
{
    const uint64_t total = (uint64_t)(m + 1) * (n + 1);
    hipLaunchKernelGGL(k_binom_table, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, binom, n, m);
    return hipGetLastError();
}

hipError_t launch_search_begin(TopState *st, uint32_t *hist, hipStream_t s)
{
    hipLaunchKernelGGL(k_search_begin, dim3(hist ? 4 : 1), dim3(256), 0, s, st, hist);
    return hipGetLastError();
}

hipError_t launch_unpack_top(const Cand *top, uint32_t count, int order, int n, const uint64_t *binom,
                             int32_t *combos, float *mi, hipStream_t s)
{
    if (!count) return hipSuccess;
    if (order > kMaxOrder) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_unpack_top, dim3((count + 255) / 256), dim3(256), 0, s, top, count, order, n, binom, combos,
                       mi);
    return hipGetLastError();
}

}  // namespace fiu

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
// Scan kernels map one wavefront to one (k-1)-SNP prefix and its 64 lanes to
// 64 consecutive candidates for the last SNP.  The prefix rows are
// wave-uniform: they are fetched with scalar loads and ANDed on the scalar
// unit, so every VALU instruction of the inner loop is either a v_and_b32
// (SGPR x VGPR) or an accumulating v_bcnt_u32_b32.  Counts stay in VGPRs;
// MI and the top-N test run in the same kernel (no count table is written).
#include "fiu_kernels.h"

#include <math.h>

namespace fiu {

template <int N> struct Pow3 { static constexpr int v = 3 * Pow3<N - 1>::v; };
template <> struct Pow3<0> { static constexpr int v = 1; };

__device__ __forceinline__ uint32_t comp(const uint4 &v, int w)
{
    return w == 0 ? v.x : w == 1 ? v.y : w == 2 ? v.z : v.w;
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

// Screening MI in FP32 from exact counts, in the conditional-entropy form
// MI = H(Y) + (1/N) sum_x [n_cx ln(n_cx/n_x) + n_tx ln(n_tx/n_x)], whose terms
// all have one sign (no cancellation inside the sum).  Error < 2e-6 nats for
// N < 2^24; only used to decide which combinations get the exact value.
template <int CELLS>
__device__ __forceinline__ float mi_screen(const uint32_t *a0, const uint32_t *a1, float hy, float inv_n)
{
    float s = 0.0f;
#pragma unroll
    for (int x = 0; x < CELLS; ++x) {
        const float fc = (float)a0[x], ft = (float)a1[x];
        const float r = __builtin_amdgcn_rcpf(fc + ft);
        const float lc = __logf(fmaxf(fc * r, 1e-30f));
        const float lt = __logf(fmaxf(ft * r, 1e-30f));
        s += fc * lc + ft * lt;
    }
    return hy + s * inv_n;
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
template <int CELLS>
__device__ __forceinline__ double mi_exact(const uint32_t *a0, const uint32_t *a1, double n, double hy)
{
#pragma clang fp contract(off)
    double hx = 0.0, hall = 0.0;
#pragma unroll
    for (int x = 0; x < CELLS; ++x) {
        const double pc = (double)a0[x] / n, pt = (double)a1[x] / n, pa = pc + pt;
        hall -= plogp(pc);
        hall -= plogp(pt);
        hx -= plogp(pa);
    }
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

// Epilogue shared by every scan kernel: screening MI, exact MI for the
// survivors, key test against the current N-th best, push.
template <int CELLS>
__device__ __forceinline__ void score_and_push(bool valid, const uint32_t *a0, const uint32_t *a1,
                                               uint64_t combo, float tau, uint64_t tau_combo,
                                               const ScanParams &P, TopState *st)
{
    bool want = false;
    float mf = 0.0f;
    if (valid) {
        const float ms = mi_screen<CELLS>(a0, a1, P.hy, P.inv_n);
        if (ms >= tau - P.delta) {
            mf = (float)mi_exact<CELLS>(a0, a1, P.n64, P.hy64);
            want = better(mf, combo, tau, tau_combo);
        }
    }
    push_candidates(want, mf, combo, P, st);
}

// ---------------------------------------------------------------------------
// Direct scan: all 3^k cells by AND + popcount (the reference's algorithm).
// ---------------------------------------------------------------------------
template <int K>
__device__ __forceinline__ void accumulate_direct(const uint4 *__restrict__ base, int np, int nq,
                                                  const int *pre, int c, uint32_t *acc)
{
    constexpr int M = K - 1, R = Pow3<M>::v;
#pragma unroll
    for (int i = 0; i < R * 3; ++i) acc[i] = 0;
    for (int q = 0; q < nq; ++q) {
        const uint4 *row = base + (size_t)q * 3 * np;
        const uint4 c0 = row[c], c1 = row[np + c], c2 = row[2 * np + c];
        uint4 pz[M][3];
#pragma unroll
        for (int i = 0; i < M; ++i)
#pragma unroll
            for (int z = 0; z < 3; ++z) pz[i][z] = row[z * np + pre[i]];
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            uint32_t pr[R];
            pr[0] = comp(pz[0][0], w);
            pr[1] = comp(pz[0][1], w);
            pr[2] = comp(pz[0][2], w);
            int size = 3;
#pragma unroll
            for (int l = 1; l < M; ++l) {
#pragma unroll
                for (int r = Pow3<M>::v / 3 - 1; r >= 0; --r) {
                    if (r < size) {
                        const uint32_t v = pr[r];
                        pr[r * 3 + 2] = v & comp(pz[l][2], w);
                        pr[r * 3 + 1] = v & comp(pz[l][1], w);
                        pr[r * 3 + 0] = v & comp(pz[l][0], w);
                    }
                }
                size *= 3;
            }
            const uint32_t x0 = comp(c0, w), x1 = comp(c1, w), x2 = comp(c2, w);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                acc[r * 3 + 0] += __builtin_popcount(pr[r] & x0);
                acc[r * 3 + 1] += __builtin_popcount(pr[r] & x1);
                acc[r * 3 + 2] += __builtin_popcount(pr[r] & x2);
            }
        }
    }
}

template <int K>
__global__ void __launch_bounds__(kBlock)
k_scan_direct(const uint4 *__restrict__ bits, const DevData D, const ScanParams P, TopState *st)
{
    constexpr int M = K - 1, CELLS = Pow3<K>::v;
    const int lane = threadIdx.x & 63;
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t j = P.j0 + (int64_t)blockIdx.x * (kBlock / kWave) + wid;
    if (j >= P.j1) return;
    const uint64_t rank = (uint64_t)P.offset + (uint64_t)j * (uint64_t)P.step;
    if (rank >= P.total_prefixes) return;
    int pre[M];
    unrank(rank, D.n, M, P.binom, pre);
    uint64_t pre_packed = 0;
#pragma unroll
    for (int i = 0; i < M; ++i) pre_packed = (pre_packed << P.combo_bits) | (uint64_t)pre[i];
    pre_packed <<= P.combo_bits;
    const int last = pre[M - 1];
    const float tau = st->tau;
    const uint64_t tau_combo = st->tau_combo;
    const uint4 *b0 = bits + D.class_off[0];
    const uint4 *b1 = bits + D.class_off[1];
    for (int t = (last + 1) >> 6; t < D.n_pad >> 6; ++t) {
        const int c = (t << 6) | lane;
        uint32_t a0[CELLS], a1[CELLS];
        accumulate_direct<K>(b0, D.n_pad, D.quads[0], pre, c, a0);
        accumulate_direct<K>(b1, D.n_pad, D.quads[1], pre, c, a1);
        score_and_push<CELLS>(c > last && c < D.n, a0, a1, pre_packed | (uint64_t)c, tau, tau_combo, P, st);
    }
}

// ---------------------------------------------------------------------------
// Top-N merge (one workgroup): current list + candidates -> best N, unique.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(1024)
k_merge(TopState *st, Cand *top, const Cand *cand, uint32_t N, uint32_t cap, uint32_t chunk_id,
        uint32_t *redo)
{
    __shared__ Cand buf[kMergeSlots];
    __shared__ uint32_t psum[1024];
    const int tid = threadIdx.x;
    const uint32_t count = min(st->cand_count, cap);
    uint32_t n_top = st->n_top;
    for (uint32_t done = 0; done < count;) {
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
        for (uint32_t k = 2; k <= L; k <<= 1) {
            for (uint32_t jj = k >> 1; jj > 0; jj >>= 1) {
                for (uint32_t i = tid; i < L; i += 1024) {
                    const uint32_t ixj = i ^ jj;
                    if (ixj > i) {
                        const Cand a = buf[i], b = buf[ixj];
                        const bool desc = (i & k) == 0;
                        const bool sw = desc ? better(b.mi, b.combo, a.mi, a.combo)
                                             : better(a.mi, a.combo, b.mi, b.combo);
                        if (sw) { buf[i] = b; buf[ixj] = a; }
                    }
                }
                __syncthreads();
            }
        }
        // unique, non-empty entries in order; contiguous segments per thread
        const uint32_t seg = (L + 1023) / 1024;
        const uint32_t s0 = min(L, tid * seg), s1 = min(L, s0 + seg);
        uint32_t local = 0;
        for (uint32_t i = s0; i < s1; ++i) {
            const bool empty = buf[i].combo == ~0ull && buf[i].mi == -INFINITY;
            const bool dup = i > 0 && buf[i].combo == buf[i - 1].combo && buf[i].mi == buf[i - 1].mi &&
                             buf[i].tag == buf[i - 1].tag;
            local += (!empty && !dup);
        }
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
            const bool empty = buf[i].combo == ~0ull && buf[i].mi == -INFINITY;
            const bool dup = i > 0 && buf[i].combo == buf[i - 1].combo && buf[i].mi == buf[i - 1].mi &&
                             buf[i].tag == buf[i - 1].tag;
            if (!empty && !dup) {
                if (pos < N) top[pos] = buf[i];
                ++pos;
            }
        }
        n_top = min(N, psum[1023]);
        __syncthreads();
        done += take;
    }
    __syncthreads();
    if (tid == 0) {
        st->n_top = n_top;
        if (n_top >= N && N > 0) {
            st->tau = top[N - 1].mi;
            st->tau_combo = top[N - 1].combo;
        } else {
            st->tau = -INFINITY;
            st->tau_combo = ~0ull;
        }
        st->cand_count = 0;
        if (st->overflow) {
            redo[st->redo_count++] = chunk_id;
            st->overflow = 0;
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
hipError_t launch_scan_direct(int order, const uint4 *bits, const DevData &D, const ScanParams &P, TopState *st,
                              hipStream_t s)
{
    const int64_t waves = P.j1 - P.j0;
    if (waves <= 0) return hipSuccess;
    const dim3 grid((unsigned)((waves + 3) / 4)), block(kBlock);
    switch (order) {
    case 2: hipLaunchKernelGGL(k_scan_direct<2>, grid, block, 0, s, bits, D, P, st); break;
    case 3: hipLaunchKernelGGL(k_scan_direct<3>, grid, block, 0, s, bits, D, P, st); break;
    case 4: hipLaunchKernelGGL(k_scan_direct<4>, grid, block, 0, s, bits, D, P, st); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_merge(TopState *st, Cand *top, const Cand *cand, uint32_t N, uint32_t cap, uint32_t chunk_id,
                        uint32_t *redo, hipStream_t s)
{
    hipLaunchKernelGGL(k_merge, dim3(1), dim3(1024), 0, s, st, top, cand, N, cap, chunk_id, redo);
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

hipError_t launch_combine_batch(const uint4 *bits, const DevData &D, const int32_t *combos, int k, uint64_t n,
                                uint64_t cw, uint64_t tw, uint64_t *out, hipStream_t s)
{
    uint64_t R = 1;
    for (int i = 0; i < k; ++i) R *= 3;
    const uint64_t total = n * R * (cw + tw);
    if (!total) return hipSuccess;
    hipLaunchKernelGGL(k_combine_batch, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, bits, D, combos, k,
                       n, cw, tw, out);
    return hipGetLastError();
}

hipError_t launch_contingency_batch(const uint4 *bits, const DevData &D, const int32_t *combos, int k, uint64_t n,
                                    uint64_t ct_size, uint32_t *out, hipStream_t s)
{
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_contingency_batch, dim3((unsigned)n), dim3(64), 0, s, bits, D, combos, k, ct_size, out);
    return hipGetLastError();
}

hipError_t launch_mi_batch(const uint32_t *tables, uint64_t n, uint64_t ct_size, double n64, double hy64, float *out,
                           hipStream_t s)
{
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_mi_batch, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, tables, n, ct_size, n64, hy64,
                       out);
    return hipGetLastError();
}

}  // namespace fiu

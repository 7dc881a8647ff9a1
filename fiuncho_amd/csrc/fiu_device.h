// Device-side data structures shared by the gfx950 kernels and the host
// runtime of fiuncho_amd.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fiu {

constexpr int kWave = 64;            // CDNA wavefront width
constexpr int kBlock = 256;          // threads per scan workgroup (4 waves)
constexpr int kMaxOrder = 12;        // orders with a scan kernel
constexpr int kMaxBatchOrder = 5;    // orders the batched engine schedules
constexpr int kMaxDirectOrder = 4;   // orders 2-4: the all-cells kernel with one prefix per wave
constexpr int kMaxKmOrder = 6;       // orders 4-6: marginal in the last SNP only
constexpr int kMinAnyOrder = 7;      // orders 7-12: the all-cells per-workgroup kernels
constexpr int kMinBigOrder = 10;     // orders 10-12: the counts in global memory
constexpr int kMarginalPW = 4;       // prefixes per wavefront in marginal scans
constexpr int kMergeSlots = 8192;    // LDS entries of the top-N merge
constexpr int kMaxOutputs = kMergeSlots / 2;   // top-N merged in LDS; above: in global memory
constexpr uint32_t kMaxOutputsBig = 1u << 24;   // single searches

// Device image of a dataset.
//
// Bits are stored "SNP-minor": for class c (0 cases, 1 controls), the 128-bit
// quad q (32-bit words 4q..4q+3 of the reference row) of genotype row z of
// SNP s lives at bits[class_off[c] + (q*3 + z)*n_pad + s].  A wavefront whose
// 64 lanes own 64 consecutive SNPs therefore reads one contiguous 1 KiB line
// per (q, z), and a wave-uniform SNP is one 16-byte scalar load.
struct DevData {
    const uint4 *bits;
    uint64_t class_off[2];   // uint4 offset of each class
    int32_t quads[2];        // 128-bit quads per row, per class
    int32_t n;               // SNPs
    int32_t n_pad;           // SNPs rounded up to a multiple of 64
    uint32_t cases, ctrls;
    const uint32_t *singles; // [n_pad][2][3] popcount of each row
    uint32_t total[2];       // samples covered per class (uniform datasets)
    int32_t uniform;         // 1 if every SNP covers the same samples
    int32_t pad_;
};

// A top-N candidate: float MI, then the combination as its lexicographic
// rank among the k-combinations of the dataset's SNPs (prefix_key), so that
// numeric order of `combo` is the lexicographic order of combinations.
struct Cand {
    float mi;
    uint32_t tag;            // dataset / request id in batched scans
    uint64_t combo;
};

// Device state of one running top-N selection.
struct TopState {
    float tau;               // MI of the current N-th best (or -inf)
    uint32_t n_top;          // entries in the top list
    uint64_t tau_combo;      // combination of the current N-th best
    uint32_t cand_count;     // candidates pushed since the last merge
    uint32_t overflow;       // a push did not fit in the buffer
    uint32_t redo_count;     // chunks that must be scanned again
    uint32_t step_overflow;  // batched scans: the last merge saw an overflow
};

struct ScanParams {
    int64_t j0, j1;          // prefix indices [j0, j1) of the distribution
    int64_t step, offset;    // Distribution step / offset (prefix lex ranks)
    uint64_t total_prefixes; // C(n, order-1)
    const uint64_t *binom;   // binom[j*(n+1)+d] = C(d, j), j <= order
    Cand *cand;              // candidate buffer
    uint32_t cand_cap;
    int32_t order;           // k (combination keys are ranks, prefix_key)
    float delta;             // screening margin on MI
    float hy;                // H(Y) in nats (from FP64)
    float inv_n;             // 1 / (cases + ctrls)
    float frac_c;            // cases / (cases + ctrls)
    float chi_scale;         // (cases + ctrls) / (cases * ctrls), 0 when a class is empty
    uint32_t bounded;        // 1: complete genotype tables, the MI upper bound applies
    uint32_t tag;
    double hy64, n64;        // H(Y) and cases+ctrls in FP64
    const uint32_t *pair;    // marginal pair table (order 3), or null
    uint32_t *hist;          // threshold sampling pass: histogram of the
                             // screening MI (kHistBins bins), else null
    uint32_t *cells;         // orders 10-12: big_slots(order) slots of
                             // 2 x 3^order counters (big_cells_bytes)
};

// Orders 10-12: resident workgroups of the persistent all-cells kernel, each
// with a slot of 2 x 3^k 32-bit counters in global memory.
__host__ __device__ inline int big_slots(int order) { return order <= 10 ? 256 : order == 11 ? 128 : 64; }
__host__ __device__ inline uint64_t big_cells_bytes(int order)
{
    uint64_t c = 1;
    for (int i = 0; i < order; ++i) c *= 3;
    return (uint64_t)big_slots(order) * 2 * c * sizeof(uint32_t);
}

// Histogram of screening MI values by the top 12 bits of their float image
// (sign, exponent, 3 mantissa bits): monotone for non-negative values, so the
// N-th best bin bounds the N-th best value from below.
constexpr int kHistBins = 4096;
__host__ __device__ inline uint32_t hist_bin(float v)
{
    union {
        float f;
        uint32_t u;
    } x;
    x.f = v > 0.0f ? v : 0.0f;
    return x.u >> 20;
}

// ---------------------------------------------------------------------------
// Batched scans: one persistent launch works through the units of many scans
// (different datasets, sizes and orders).  A unit is one wavefront's work: a
// group of prefixes (kMarginalPW for marginal kinds, 1 for direct kinds)
// against a chunk of `tiles` consecutive tiles of 64 last-SNP candidates.
// ---------------------------------------------------------------------------
enum BatchKind : int32_t {
    kKindScan2m = 0,   // order 2, marginal
    kKindScan3m = 1,   // order 3, marginal (pair table in P.pair)
    kKindDirect2 = 2,  // order 2, all cells
    kKindDirect3 = 3,  // order 3, all cells
    kKindDirect4 = 4,  // order 4, all cells (its own launch: register budget)
    kKindMfma2 = 5,    // order 2 on the matrix cores (fp4): units are 64 x 64 tiles of super-blocks
    kKindMfma3 = 6,    // order 3 on the matrix cores (fp4): units are (first SNP, tile) pairs
    kKindKm4 = 7,      // order 4, marginal in the last SNP (launch 3)
    kKindKm5 = 8,      // order 5, marginal in the last SNP (launch 3)
    kKindKm5All = 9,   // order 5, every cell counted (any dataset; launch 3)
};

// fp4 image of a dataset's genotype rows 0 and 1 (operands of the matrix-core
// scans): the 128 samples of quad q, genotype g, SNP s of a class are 64
// bytes = 4 uint4 at class_off[c] + 4 ((2 q + g) n_pad + s).
struct Fp4Image {
    const uint4 *base;
    uint64_t class_off[2];   // uint4 offset of each class
};

__host__ __device__ inline int batch_group(int32_t kind) { return kind <= kKindScan3m ? kMarginalPW : 1; }
__host__ __device__ inline bool batch_mfma(int32_t kind) { return kind == kKindMfma2 || kind == kKindMfma3; }

struct BatchJob {
    DevData D;
    ScanParams P;            // distribution, MI constants, candidate buffer
    TopState *st;
    Cand *top;
    uint32_t outputs;
    int32_t kind;            // BatchKind
    int32_t chunks;          // tile chunks per prefix group
    int32_t tiles;           // tiles per chunk
    Fp4Image F;              // matrix-core kinds: the dataset's fp4 image
    int32_t T;               // matrix-core kinds: 64-SNP tiles (n_pad / 64)
    int32_t wide;            // matrix-core kinds: the 64 x 64 (order 3) / 64 x 128 (order 2) tiles
};

// A contiguous prefix range [j0, j1) of one job in one launch; its units are
// numbered from `start` (prefix group major, tile chunk minor).  Matrix-core
// kinds: [j0, j1) are first SNPs (order 3) or super-blocks (order 2), and
// `start` numbers the units of each XCD's queue (every XCD has as many).
struct BatchSeg {
    uint64_t start;
    int64_t j0, j1;
    int32_t job;
    int32_t pad_;
};

// Per-launch header in device memory (the launch itself is a fixed graph).
struct BatchHeader {
    uint32_t nseg[3];        // segments of the main launch, the direct order-4
                             // launch, the marginal order-4/5 launch
    uint32_t units[3];       // units of each launch
    uint32_t nmerge;         // jobs to merge after the launches
    uint32_t head[3];        // work counters (cleared by the graph)
    uint32_t mseg;           // segments of the matrix-core launch
    uint32_t munits;         // units of each XCD's queue in it
    uint32_t mhead[8];       // its per-XCD work counters (cleared by the graph)
};

__host__ __device__ inline uint64_t pair_index(int64_t b, int64_t c, int64_t n)
{
    // row-major strict upper triangle: pairs (b, c), b < c < n
    return (uint64_t)(b * (2 * n - b - 1) / 2 + (c - b - 1));
}

}  // namespace fiu

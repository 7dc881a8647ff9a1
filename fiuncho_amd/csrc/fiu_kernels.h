// Host-side launchers of the gfx950 kernels (defined in fiu_kernels.hip).
#pragma once

#include "fiu_device.h"

namespace fiu {

// Orders 4 to 6 with the prefix rows split over lanes: marginal in the last
// SNP (uniform datasets), or all_cells: every cell counted (any dataset);
// P.hist set: the threshold sampling pass.
hipError_t launch_scan_km(int order, const DevData &D, const ScanParams &P, TopState *st, hipStream_t s,
                          bool all_cells = false);
hipError_t launch_scan_direct(int order, const DevData &D, const ScanParams &P, TopState *st, hipStream_t s);
// pw: prefixes per wavefront (2, 4 or, order 2 only, 8; the threshold
// sampling pass uses 4)
hipError_t launch_scan_marginal(int order, const DevData &D, const ScanParams &P, TopState *st, hipStream_t s,
                                int pw = kMarginalPW);
hipError_t launch_tau_from_hist(const uint32_t *hist, uint32_t N, float delta, TopState *st, hipStream_t s);
// A search's start: state reset, sampling histogram (when set) cleared.
hipError_t launch_search_begin(TopState *st, uint32_t *hist, hipStream_t s);
hipError_t launch_pair_table(const DevData &D, const ScanParams &P, uint32_t *pair_out, hipStream_t s);
// Order 2 on the matrix cores: the pair triangle of T x T 64-SNP tiles is cut
// into super-blocks of 8 x 8 tiles; a launch scores super-blocks [s0, s1)
// into the top-N state, with fp4 (e2m1) operands from the dataset's fp4 image
// or, when `fp4` is null, int8 operands expanded from the bit tables.
int64_t mfma2_superblocks(int T);
// fp4 image of genotype rows 0 and 1 (operands of the matrix-core scans):
// uint4 units per class, and the kernel that writes both classes (class 1
// after class 0).
uint64_t fp4_image_units(const DevData &D, int cls);
hipError_t launch_fp4_image(const DevData &D, uint4 *out, hipStream_t s);
// Order 3 on the matrix cores (fp4): first SNPs [a0, a1), every (b, c) past
// them, with the pair table in P.pair and the dataset's fp4 image; `ring`:
// 3 or 4 LDS stages of 64 x 32 (b, c) tiles, 5 or 6 (3 or 4 stages) of 64 x
// 64 tiles (classes below 2^16 samples).
hipError_t launch_mfma3(const DevData &D, const ScanParams &P, TopState *st, const uint4 *fp4, int a0, int a1, int T,
                        int ring, hipStream_t s);
// Order 2 on the matrix cores: super-blocks [s0, s1) of 8 x 8 tiles, fp4
// operands from the image `fp4` (int8 when null); `tile`: pairs per
// workgroup, 0 = 64 x 64, 1 = 64 x 128, 2 = 128 x 128 (1 and 2: fp4, classes
// below 2^16 samples).
hipError_t launch_mfma2(const DevData &D, const ScanParams &P, TopState *st, int64_t s0, int64_t s1, int T,
                        const uint4 *fp4, int tile, hipStream_t s);
hipError_t launch_merge(TopState *st, Cand *top, const Cand *cand, uint32_t N, uint32_t cap, uint32_t chunk_id,
                        uint32_t *redo, hipStream_t s);
// Top-N merge for N above kMaxOutputs, in global memory: the list's n_top
// entries and `count` candidates (sizes read by the host) sorted in `work`
// (L entries, a power of two >= n_top + count), the first N unique kept.
hipError_t launch_merge_big(TopState *st, Cand *top, uint32_t n_top, const Cand *cand, uint32_t count, uint32_t N,
                            Cand *work, uint32_t L, uint32_t chunk_id, uint32_t *redo, hipStream_t s);
// Genotypes (SNP-major int8; the buffer allocated to a multiple of 16 bytes)
// -> reference-layout words of snps x 3(cw+tw).  `idx` null: class c is the
// run of n_class[c] samples from start[c]; else the per-class sample lists.
// `zero`: per word slot, 1 = written as 0 (or null).
constexpr size_t kPackLdsMax = 64 * 1024;
inline size_t pack_lds_bytes(uint64_t samples, uint64_t cw, uint64_t tw)
{
    return (size_t)(cw + tw) * 24 + (size_t)((samples + 31) / 16 * 16);
}
hipError_t launch_pack_words(const int8_t *geno, uint64_t snps, uint64_t samples, const int32_t *idx,
                             const uint32_t n_class[2], const uint32_t start[2], uint64_t cw, uint64_t tw,
                             const uint8_t *zero, uint64_t *words, hipStream_t s);
hipError_t launch_transpose(const uint64_t *src, uint4 *dst, int n, int n_pad, int cls, uint64_t cw, uint64_t tw,
                            int quads, uint64_t class_off, hipStream_t s);
hipError_t launch_singles(const uint4 *bits, const DevData &D, uint32_t *singles, uint32_t *flags, hipStream_t s);
hipError_t launch_combine_batch(const DevData &D, const int32_t *combos, int k, uint64_t n, uint64_t cw, uint64_t tw,
                                uint64_t *out, hipStream_t s);
hipError_t launch_contingency_batch(const DevData &D, const int32_t *combos, int k, uint64_t n, uint64_t ct_size,
                                    uint32_t *out, hipStream_t s);
hipError_t launch_mi_batch(const uint32_t *tables, uint64_t n, uint64_t ct_size, double n64, double hy64, float *out,
                           hipStream_t s);

// Batched scans: resident workgroups of the persistent scan kernel of launch
// 0 (popcount orders 2-3), 1 (direct order 4), 2 (matrix-core kinds, a
// multiple of 8) or 3 (marginal orders 4-5) on the current device, and the
// whole batched step: counter reset, the four persistent scans, the per-job
// merges.
int batch_blocks(int launch);
hipError_t launch_batch(const BatchJob *jobs, const BatchSeg *segs0, const BatchSeg *segs1, const BatchSeg *segs2,
                        const BatchSeg *segs3, const int32_t *merge_list, BatchHeader *hdr, int blocks0, int blocks1,
                        int blocks2, int blocks3, int merge_blocks, hipStream_t s);

// Searches in a fixed device block: the binomial table C(x, j), x <= n,
// j <= m (saturated), and the top list unpacked to ascending SNP indices
// (binom up to row `order`) and float MI.
hipError_t launch_binom_table(uint64_t *binom, uint64_t n, int m, hipStream_t s);
hipError_t launch_unpack_top(const Cand *top, uint32_t count, int order, int n, const uint64_t *binom,
                             int32_t *combos, float *mi, hipStream_t s);

}  // namespace fiu

/*
 * CPU oracle for the exhaustive epistasis search.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls
 * this code; only tests/, the smoke check and the CPU-baseline leg of the
 * benchmark load it, and only as the checker (or as the timed CPU baseline).
 *
 * It restates, in plain C, the algorithm of the reference search path
 * (UDC-GAC/fiuncho).  Every function names the reference file:line it follows.
 * Parity of each function is pinned by tests/test_oracle.py against the
 * reference's own fixtures (tests/data, tests/mi.cpp known answers, the
 * MPI3SNP golden list in tests/utils.h) and against vectors written by the
 * reference compiled from source (oracle/_ref).
 */
#ifndef FIU_ORACLE_H
#define FIU_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A loaded dataset in the reference's host layout (GenotypeTable.h:77-101):
 * SNP s occupies words [s*3*(cw+tw), (s+1)*3*(cw+tw)): three case rows of cw
 * words, then three control rows of tw words. */
typedef struct {
    size_t snps, cases, ctrls, cases_words, ctrls_words;
    uint64_t *words;
} orc_dataset;

/* Words per class row for n samples, padded to `align` bytes like
 * RAWFile.hpp:299-309 / TPEDFile.hpp:282-292 (align = 0: no padding). */
size_t orc_words_for(size_t n, int align);

/* Loaders.  Return 0 on success, -1 on error with a message in err. */
int orc_read_raw(const char *path, int align, orc_dataset *out, char *err,
                 size_t errlen);
int orc_read_tped(const char *tped, const char *tfam, int align,
                  orc_dataset *out, char *err, size_t errlen);
void orc_dataset_free(orc_dataset *d);

/* Pack a genotype matrix (snps x samples, row-major, values 0/1/2; any other
 * value sets no bit) with per-sample control flags into the layout above. */
int orc_pack(const int8_t *geno, const uint8_t *is_ctrl, size_t snps,
             size_t samples, int align, orc_dataset *out);

/* GenotypeTable::combine semantics (src/base/gt_combine.cpp:3-18). */
void orc_combine(const uint64_t *t1, size_t size1, const uint64_t *t2,
                 size_t words, uint64_t *out);
/* combine_and_popcnt semantics (src/base/gt_popcnt.cpp:4-23). */
void orc_popcnt(const uint64_t *t1, size_t size1, const uint64_t *t2,
                size_t words, uint32_t *ct, size_t ct_size);

/* Padded contingency-table length per class (ContingencyTable.h:53-65). */
size_t orc_ct_size(int order, int align);

/* Counts of one combination (ascending SNP indices) through the prefix chain
 * of ThreadedSearch.h:160-176.  cases_ct / ctrls_ct hold 3^k entries. */
void orc_counts(const orc_dataset *d, const int *combo, int k,
                uint32_t *cases_ct, uint32_t *ctrls_ct);

/* Brute-force counts straight from the genotype matrix (closed form). */
void orc_counts_bruteforce(const int8_t *geno, const uint8_t *is_ctrl,
                           size_t samples, const int *combo, int k,
                           uint32_t *cases_ct, uint32_t *ctrls_ct);

/* Mutual information of a contingency table of `size` cells per class.
 *  _ref512: FP32, the operation order of src/avx512bw/mi_compute_512.cpp:20-69
 *           (needs AVX-512F at run time and glibc libmvec);
 *  _base:   FP32, the order of src/base/mi_compute.cpp:14-41;
 *  _f64:    FP64 closed form from the integer counts. */
float orc_mi_f32_ref512(const uint32_t *cases_ct, const uint32_t *ctrls_ct,
                        size_t size, unsigned cases, unsigned ctrls);
float orc_mi_f32_base(const uint32_t *cases_ct, const uint32_t *ctrls_ct,
                      size_t size, unsigned cases, unsigned ctrls);
double orc_mi_f64(const uint32_t *cases_ct, const uint32_t *ctrls_ct,
                  size_t size, unsigned cases, unsigned ctrls);

/* Distribution<int> (Distribution.h:54-130): write the (k-1)-prefixes of the
 * distribution (n, k1, step, offset) to out (k1 ints each), at most max
 * prefixes.  Returns the number of prefixes in the distribution. */
size_t orc_dist_prefixes(int n, int k1, int step, int offset, int *out,
                         size_t max);

/* Exhaustive search (ThreadedSearch.h:142-182, 207-246).
 *  mi_mode 0: FP32 in the AVX-512 order, 1: FP32 base order, 2: FP64 rounded
 *  to float.  Per-thread top-N keeps first-seen on equal value (MaxArray.h:
 *  31-46); the merged list is ordered by (value desc, combination lex asc).
 *  Writes up to `outputs` results; returns the count, or -1 on error. */
long orc_search(const orc_dataset *d, int order, int step, int offset,
                unsigned outputs, int threads, int mi_mode, int *combos,
                float *mi, uint64_t *evaluated);

/* 1 if this CPU can run the AVX-512 paths. */
int orc_has_avx512(void);

#ifdef __cplusplus
}
#endif
#endif

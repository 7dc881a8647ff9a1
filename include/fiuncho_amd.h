/*
 * fiuncho_amd -- C ABI of the MI355X (gfx950) exhaustive epistasis search.
 *
 * The boundary mirrors the reference's search path (UDC-GAC/fiuncho):
 *   - a dataset is the bit-packed case/control genotype tables exactly as
 *     Dataset<uint64_t>::read leaves them in host memory
 *     (include/fiuncho/Dataset.h:48-94, GenotypeTable.h:77-101);
 *   - fiu_search is Search::run (include/fiuncho/Search.h:40-43), the call
 *     ThreadedSearch::run makes on the CPU (ThreadedSearch.h:207-246);
 *   - the batch operators are GenotypeTable::combine / combine_and_popcnt
 *     (GenotypeTable.h:123-140) and MutualInformation<float>::compute
 *     (algorithms/MutualInformation.h:83-84) applied to many tables per call.
 *
 * Conventions: every entry point returns a fiu_status; no C++ exception
 * crosses the boundary; fiu_last_error() holds a thread-local message for the
 * last failing call on the calling thread.  A dataset handle may be used by
 * one thread at a time; distinct handles are independent.  All pointers are
 * host pointers unless stated otherwise.
 */
#ifndef FIUNCHO_AMD_H
#define FIUNCHO_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    FIU_OK = 0,
    FIU_ERR_INVALID = 1,     /* bad argument (the reference throws or UB) */
    FIU_ERR_DEVICE = 2,      /* HIP runtime or kernel error */
    FIU_ERR_NOMEM = 3,       /* device or host allocation failed */
    FIU_ERR_UNSUPPORTED = 4, /* valid input outside this build's limits */
    FIU_ERR_IO = 5,          /* file could not be read or parsed */
    FIU_ERR_REJECTED = 6,    /* scan server: admission refused */
    FIU_ERR_CANCELLED = 7,   /* scan server: request cancelled */
    FIU_ERR_PENDING = 8      /* scan server: result not ready yet */
} fiu_status;

/* Thread-local message of the last failing call on this thread ("" if none). */
const char *fiu_last_error(void);
/* Library version string. */
const char *fiu_version(void);
/* Select the HIP device used by handles created afterwards on this thread. */
fiu_status fiu_set_device(int device);

/* ------------------------------------------------------------------------ */
/* Datasets                                                                  */
/* ------------------------------------------------------------------------ */

typedef struct fiu_dataset fiu_dataset;

/* Upload a dataset.  `words` is the reference host layout: SNP s occupies
 * 3*(cases_words+ctrls_words) words starting at s*3*(cases_words+ctrls_words):
 * three case rows (genotype 0,1,2) of cases_words words, then three control
 * rows of ctrls_words words (GenotypeTable.h:77-101, RAWFile.hpp:209-290).
 * The words are copied; the caller keeps ownership of `words`.
 * Replaces: the Dataset<uint64_t> that Search::run borrows (Search.h:40-43,
 * Dataset.h:123-151). */
fiu_status fiu_dataset_create(const uint64_t *words, size_t snps,
                              size_t cases_words, size_t ctrls_words,
                              uint32_t cases, uint32_t ctrls,
                              fiu_dataset **out);

/* Read .raw, or .tped + .tfam (in either order), pack the bit tables with the
 * layout of RawFile::fill / TPEDFile::fill (ALIGN = 64 bytes) and upload.
 * Replaces: Dataset<uint64_t>::read (Dataset.h:48-94). */
fiu_status fiu_dataset_read(const char *const *paths, size_t npaths,
                            fiu_dataset **out);

/* Host-only read of the same files into the reference host layout, without a
 * device: call with words = NULL to learn the shape, then with a buffer of
 * at least snps*3*(cases_words+ctrls_words) words (`capacity`).
 * Replaces: the tables Dataset<uint64_t>::read builds (Dataset.h:48-94,
 * RAWFile.hpp:292-313, TPEDFile.hpp:271-296). */
fiu_status fiu_read_files(const char *const *paths, size_t npaths,
                          uint64_t *words, size_t capacity, size_t *snps,
                          uint32_t *cases, uint32_t *ctrls,
                          size_t *cases_words, size_t *ctrls_words);

/* Dataset shape (any out pointer may be NULL). */
fiu_status fiu_dataset_info(const fiu_dataset *ds, size_t *snps,
                            uint32_t *cases, uint32_t *ctrls,
                            size_t *cases_words, size_t *ctrls_words,
                            size_t *device_bytes);

/* Copy the host-layout words back (3*(cases_words+ctrls_words)*snps words). */
fiu_status fiu_dataset_words(const fiu_dataset *ds, uint64_t *out);

fiu_status fiu_dataset_destroy(fiu_dataset *ds);

/* ------------------------------------------------------------------------ */
/* Search                                                                    */
/* ------------------------------------------------------------------------ */

typedef enum {
    FIU_ALGO_AUTO = 0,     /* fastest exact path the dataset allows */
    FIU_ALGO_DIRECT = 1,   /* AND+popcount of all 3^k cells (the reference's) */
    FIU_ALGO_MARGINAL = 2  /* 2^k cells + inclusion-exclusion from exact
                              lower-order counts; needs every SNP to cover the
                              same samples (checked at upload) */
} fiu_algorithm;

/* fiu_config.flags */
#define FIU_FLAG_NO_THRESHOLD_SAMPLE 1u  /* skip the pass that seeds the top-N
                                            threshold from a sample */

typedef struct {
    uint32_t struct_size;    /* sizeof(fiu_config) */
    int32_t algorithm;       /* fiu_algorithm */
    uint64_t chunk_combos;   /* combinations per scan launch, 0 = auto */
    uint32_t candidate_cap;  /* device candidate buffer entries, 0 = auto */
    uint32_t flags;          /* FIU_FLAG_* */
    void *stream;            /* hipStream_t to run on, NULL = library stream */
} fiu_config;

/* Fill a fiu_config with defaults. */
void fiu_config_init(fiu_config *cfg);

typedef struct {
    uint64_t combinations;   /* combinations evaluated */
    uint64_t launches;       /* scan kernel launches */
    uint64_t redo_chunks;    /* chunks re-scanned after a candidate overflow */
    double scan_ms;          /* device time of the scan launches (HIP events) */
    double total_ms;         /* wall time of the call */
    int32_t algorithm;       /* algorithm actually used */
    int32_t reserved;
} fiu_search_stats;

/* Exhaustive order-k search over the combinations whose (k-1)-prefix belongs
 * to Distribution<int>(snps, order-1, dist_step, dist_offset)
 * (Distribution.h:147-176), keeping the `outputs` best by mutual information.
 * Results are ordered by MI descending, ties by combination ascending;
 * out_combos receives order ints per result (ascending SNP indices), out_mi
 * the MI as float, *out_count the number of results (<= outputs).
 * Replaces: Search::run (Search.h:40-43) / ThreadedSearch::run
 * (ThreadedSearch.h:207-246).  `cfg` and `stats` may be NULL. */
fiu_status fiu_search(const fiu_dataset *ds, uint16_t order,
                      int32_t dist_step, int32_t dist_offset,
                      uint32_t outputs, const fiu_config *cfg,
                      int32_t *out_combos, float *out_mi, uint32_t *out_count,
                      fiu_search_stats *stats);

/* ------------------------------------------------------------------------ */
/* Batch operators (one GPU launch for many tables)                          */
/* ------------------------------------------------------------------------ */

/* Padded contingency-table length per class for `order`
 * (ContingencyTable.h:53-65 with ALIGN = 64). */
size_t fiu_ct_size(uint16_t order);

/* For each of n ascending combinations of `order` SNPs: the order-`order`
 * genotype table built by the chain of GenotypeTable::combine calls
 * (GenotypeTable.h:123-124, ThreadedSearch.h:160-164).  out receives, per
 * combination, 3^order case rows of cases_words words then 3^order control
 * rows of ctrls_words words (the GenotypeTable layout). */
fiu_status fiu_combine_batch(const fiu_dataset *ds, const int32_t *combos,
                             uint16_t order, size_t n, uint64_t *out);

/* For each of n ascending combinations: the contingency table that
 * GenotypeTable::combine_and_popcnt leaves (GenotypeTable.h:137-140):
 * out_counts gets fiu_ct_size(order) case counts then as many control counts
 * (cells past 3^order are 0); out_mi (may be NULL) gets the mutual
 * information of each table (MutualInformation<float>::compute). */
fiu_status fiu_contingency_batch(const fiu_dataset *ds, const int32_t *combos,
                                 uint16_t order, size_t n,
                                 uint32_t *out_counts, float *out_mi);

/* Mutual information of n caller-given tables, each fiu_ct_size(order) case
 * counts then as many control counts, for a dataset of `cases` cases and
 * `ctrls` controls (MutualInformation.h:58-65, 83-84). */
fiu_status fiu_mi_batch(const uint32_t *tables, size_t n, uint16_t order,
                        uint32_t cases, uint32_t ctrls, float *out_mi);


/* ------------------------------------------------------------------------ */
/* Synthetic data (no counterpart in the reference: its tests read files)    */
/* ------------------------------------------------------------------------ */

/* Hardy-Weinberg genotypes with per-SNP minor-allele frequency
 * MAF ~ U(maf_lo, maf_hi) (or, with probability low_frac, U(low_lo, low_hi));
 * samples 0..cases-1 are cases, the rest controls; optionally one planted
 * combination whose joint genotype follows a random penetrance table
 * U(pen_lo, pen_hi) drawn from penetrance_seed.  Every draw is a hash of
 * (seed, stream, counter), so the data is identical on any machine. */
typedef struct {
    uint32_t struct_size;     /* sizeof(fiu_synth_spec) */
    uint32_t cases, ctrls;
    uint32_t planted_order;   /* 0 = nothing planted */
    uint64_t snps;
    double maf_lo, maf_hi;
    double low_frac, low_lo, low_hi;
    uint64_t seed;
    int64_t planted[8];       /* ascending SNP indices of the planted combination */
    uint64_t penetrance_seed;
    double pen_lo, pen_hi;
    int32_t threads;          /* host threads, 0 = all */
    int32_t reserved;
} fiu_synth_spec;

void fiu_synth_spec_init(fiu_synth_spec *spec);

/* Generate into host memory: `words` (reference layout, may be NULL) and/or
 * `geno` (snps x samples int8 genotypes, may be NULL).  Row lengths follow
 * the ALIGN = 64 rule and are returned in *cases_words / *ctrls_words. */
fiu_status fiu_synth_generate(const fiu_synth_spec *spec, uint64_t *words,
                              int8_t *geno, size_t *cases_words,
                              size_t *ctrls_words);

/* Generate and upload in one call. */
fiu_status fiu_dataset_synth(const fiu_synth_spec *spec, fiu_dataset **out);

#ifdef __cplusplus
}
#endif
#endif

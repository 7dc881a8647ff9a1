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

/* Read .raw, or .tped + .tfam (in either order), parse them on the host and
 * pack the bit tables on the device with the layout of RawFile::fill /
 * TPEDFile::fill (ALIGN = 64 bytes), as fiu_dataset_from_genotypes with
 * FIU_GENO_READER_FLUSH.
 * Replaces: Dataset<uint64_t>::read (Dataset.h:48-94). */
fiu_status fiu_dataset_read(const char *const *paths, size_t npaths,
                            fiu_dataset **out);

/* Pack a genotype matrix on the device and keep it there.  `geno` is
 * SNP-major, snps x samples int8: 0, 1, 2 = minor-allele count, any other
 * value = uncalled (sets no bit).  `is_ctrl[j]` non-zero makes sample j a
 * control, zero a case.  The tables are the ones RawFile::fill and
 * TPEDFile::fill build from these genotypes (RAWFile.hpp:209-290,
 * TPEDFile.hpp:175-269, ALIGN = 64 bytes): samples keep their order within
 * their class, full words hold them most-significant bit first, the last
 * partial word is right-aligned, padding bits and words are zero.  With
 * FIU_GENO_READER_FLUSH the words that the reference readers write a second
 * time from an emptied buffer (RAWFile.hpp:226-259, TPEDFile.hpp:213-245)
 * are zero too, which makes the tables equal to what Dataset<uint64_t>::read
 * builds from a file holding these genotypes; fiu_dataset_read packs this way.
 * Replaces: the packing half of Dataset<uint64_t>::read (Dataset.h:48-94). */
#define FIU_GENO_READER_FLUSH 1u
fiu_status fiu_dataset_from_genotypes(const int8_t *geno, size_t snps,
                                      size_t samples, const uint8_t *is_ctrl,
                                      uint32_t flags, fiu_dataset **out);

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
    FIU_ALGO_DIRECT = 1,   /* AND+popcount of all 3^k cells (the reference's);
                              every order, any dataset (the only algorithm
                              at orders 7-12) */
    FIU_ALGO_MARGINAL = 2, /* orders 2-3: 2^k cells + inclusion-exclusion
                              from exact lower-order counts; orders 4-6: the
                              prefix's 3^(k-1) rows against genotypes 0 and 1
                              of the last SNP, genotype 2 from the prefix
                              rows' counts; needs every SNP to cover the same
                              samples (checked at upload) */
    FIU_ALGO_MFMA = 3      /* orders 2 and 3: the 2^k marginal cells as
                              indicator products on the matrix cores (fp4
                              with fp32 accumulation, or int8 with int32;
                              exact); same requirement, and for order 3 the
                              pair table */
} fiu_algorithm;

/* fiu_config.flags */
#define FIU_FLAG_NO_THRESHOLD_SAMPLE 1u  /* skip the pass that seeds the top-N
                                            threshold from a sample */
#define FIU_FLAG_MFMA_INT8 2u            /* FIU_ALGO_MFMA: int8 operands with
                                            int32 accumulation instead of the
                                            default fp4 (e2m1) operands with
                                            fp32 accumulation (exact below
                                            2^24 samples per class) */

typedef struct {
    uint32_t struct_size;    /* sizeof(fiu_config) */
    int32_t algorithm;       /* fiu_algorithm */
    uint64_t chunk_combos;   /* combinations per scan launch, 0 = auto */
    uint32_t candidate_cap;  /* device candidate buffer entries, 0 = auto */
    uint32_t flags;          /* FIU_FLAG_* */
    void *stream;            /* hipStream_t to run on, NULL = library stream */
    int32_t variant;         /* kernel variant of the algorithm, 0 = the
                                tuned or default one (fiu_tune_entry) */
    int32_t reserved;
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
    int32_t variant;         /* its kernel variant (fiu_tune_entry) */
} fiu_search_stats;

/* Exhaustive order-k search (k = 2..12, any dataset; a result is keyed by
 * its lexicographic rank among the C(snps, k) combinations, so C(snps, k)
 * must stay below 2^64 - 1: at most 145,056 SNPs at order 4, 18,580 at
 * order 5, 4,868 at order 6, 1,913 at order 7, 967 at order 8, 577 at order
 * 9, 386 at order 10, 282 at order 11, 218 at order 12) over the combinations whose (k-1)-prefix belongs
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

/* Search that allocates nothing.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  Same arguments, results, errors and algorithm choice as
 * fiu_search; the search runs in one fixed block of device memory per GPU
 * (fiu_search_static_bytes(), under 9 MiB), which the library holds until
 * the process ends, and calls no hipMalloc / hipFree (and no host heap
 * allocation).  Limits, refused with FIU_ERR_UNSUPPORTED: candidate_cap <=
 * 2^18 (default 2^18); (order + 1) * (snps + 1) <= 2^19 (binomial table),
 * e.g. 131,071 SNPs at order 3, 52,427 at order 9; orders up to 9; outputs
 * up to 4,096; 2^14 scan launches (chunks grow to fit).  The dataset's pair
 * table and fp4 image are used when an earlier fiu_search built them;
 * otherwise the kernels that need neither run
 * (same results).  The matrix-core scans run their tiles that keep no counts
 * in per-lane private memory (order 3: variants 3 / 4, order 2: the 64 x 64
 * tile), so no launch needs scratch memory from the runtime (same results).
 * stats->scan_ms is 0 (no timing events).  Every call on one
 * GPU works in that GPU's block, which keeps the last call's contents (each
 * call clears what must start at zero): calls that overlap on one GPU race,
 * and must be serialised by the caller. */
fiu_status fiu_search_static(const fiu_dataset *ds, uint16_t order,
                             int32_t dist_step, int32_t dist_offset,
                             uint32_t outputs, const fiu_config *cfg,
                             int32_t *out_combos, float *out_mi,
                             uint32_t *out_count, fiu_search_stats *stats);
/* Bytes of the fixed block fiu_search_static uses on each GPU. */
size_t fiu_search_static_bytes(void);

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
/* Result formats (host only)                                                */
/* ------------------------------------------------------------------------ */

/* The text of one result as Result<int, float>::str writes it
 * (include/fiuncho/utils/Result.h:37-45): each SNP index followed by a
 * space, then the MI as a default-formatted stream writes a float (6
 * significant digits).  Writes at most cap - 1 characters and a NUL (buf may
 * be NULL with cap 0); returns the full length, as snprintf does.
 * Replaces: Result::str. */
size_t fiu_result_str(const int32_t *combo, uint16_t order, float mi,
                      char *buf, size_t cap);

/* The binary record of Result<int, float>::serialize (Result.h:47-55):
 * the order as a size_t, the order indices as int32, the MI as float, in host
 * byte order.  Returns the record's size; writes it only when cap is at least
 * that (buf may be NULL to ask for the size).
 * Replaces: Result::serialize. */
size_t fiu_result_serialize(const int32_t *combo, uint16_t order, float mi,
                            void *buf, size_t cap);

/* Read one record written by Result::serialize / fiu_result_serialize
 * (Result.h:57-66) from buf (len bytes available): *order indices into combo
 * (room for max_order), the MI into *mi, the record's size into *used (may be
 * NULL).  FIU_ERR_INVALID for a truncated record or an order above
 * max_order.
 * Replaces: Result::deserialize. */
fiu_status fiu_result_deserialize(const void *buf, size_t len, int32_t *combo,
                                  uint16_t max_order, uint16_t *order,
                                  float *mi, size_t *used);

/* Write count results (order indices each in combos, MI in mi) to a text
 * file, one fiu_result_str line and a newline each, as the reference's
 * command line writes its output file (app/main.cpp:152-158); FIU_ERR_IO when
 * the file cannot be written. */
fiu_status fiu_write_results(const char *path, const int32_t *combos,
                             const float *mi, uint32_t count, uint16_t order);

/* ------------------------------------------------------------------------ */
/* Tuning table                                                              */
/* ------------------------------------------------------------------------ */
/* No counterpart in the reference (one CPU kernel per build).  AUTO searches
 * and batched scans whose (order, SNPs, cases + ctrls) fall in a class of the
 * table use that class's algorithm and kernel variant, when the dataset
 * allows them (else the default choice).  Variants:
 *   FIU_ALGO_MARGINAL: prefixes per wavefront, 2, 4 (default) or, order 2
 *                      only, 8;
 *   FIU_ALGO_MFMA, order 3: tile and LDS staging depth: 3, 4 = 64 x 32
 *                      (b, c) per workgroup with 3 or 4 stages, 5, 6 =
 *                      64 x 64 with 3 or 4 stages (default: 6 from 2,048
 *                      SNPs and 8,192 samples on, else 4; 5 and 6 fall back
 *                      to 4 when a class holds 2^16 samples or more);
 *   FIU_ALGO_MFMA, order 2: 1 = int8 operands, 2 = fp4 operands, 64 x 64
 *                      pairs per workgroup, 3 = fp4 operands, 64 x 128 pairs
 *                      per workgroup, 4 = fp4 operands, 128 x 128 pairs per
 *                      workgroup (3 and 4: 64 x 64 when a class holds 2^16
 *                      samples or more); default: 4 from 16,384 SNPs and
 *                      8,192 samples on, else 2.
 * The table is written by an autotuner that times the candidates on the
 * device (fiuncho_amd.tune); it is process-wide and starts empty.  Batched
 * scans use its algorithm and its batch_unit.  For the popcount kinds
 * (orders 2-5 off the matrix cores) batch_unit is the work of one unit of the
 * persistent kernels, 0 = the default (about 40,000 wave instructions: that
 * many 64-SNP tiles of last SNPs per prefix group), k = 1..6 = the default
 * times 2^(k-3) (a quarter to 8 times).  For FIU_ALGO_MFMA entries it is the
 * batched tile: 0 = by size (64 x 64 / 64 x 128 from 2,048 / 16,384 SNPs and
 * 8,192 samples on, as the single search's default), 1 = 64 x 32 (order 3)
 * / 64 x 64 (order 2), 2 = 64 x 64 / 64 x 128 (classes below 2^16 samples);
 * above 2 is FIU_ERR_INVALID.  The variant does not steer batched scans. */
typedef struct {
    uint16_t order;
    uint16_t batch_unit;              /* batched kinds' unit size (popcount) or tile (MFMA) */
    uint32_t snps_lo, snps_hi;        /* SNPs in [lo, hi) */
    uint32_t samples_lo, samples_hi;  /* cases + ctrls in [lo, hi) */
    int32_t algorithm;                /* fiu_algorithm (not AUTO) */
    int32_t variant;                  /* 0 = the algorithm's default */
    float ms;                         /* time measured by the tuner (information) */
} fiu_tune_entry;

/* Replace the table (n = 0 clears it); FIU_ERR_INVALID for an unknown
 * algorithm / variant or an empty range, and the table is left unchanged. */
fiu_status fiu_set_tuning(const fiu_tune_entry *entries, size_t n);
/* Copy up to cap entries; *n = entries in the table. */
fiu_status fiu_get_tuning(fiu_tune_entry *entries, size_t cap, size_t *n);

/* ------------------------------------------------------------------------ */
/* Batched scans and the scan server                                         */
/* ------------------------------------------------------------------------ */
/* No counterpart in the reference, which runs one search per process
 * (MPIEngine.h:135-137).  Many searches over different datasets, sizes and
 * orders (2-5) advance together through batched device steps: persistent launches
 * over a slice of every active search (popcount kinds, orders 4-5, matrix-core
 * kinds for order 2 and whole order-3 scans), then one top-N merge per
 * search, replayed from a HIP graph.  Each search returns exactly what
 * fiu_search returns for it alone. */

typedef struct {
    uint32_t struct_size;    /* sizeof(fiu_batch_config) */
    uint32_t max_active;     /* searches on the device at once (default 128,
                                at most 1024) */
    uint64_t memory_budget;  /* server: device bytes for the tables it
                                uploads, 0 = a quarter of the free
                                memory, at most 8 GiB */
    double step_ms;          /* target device time of one batched step;
                                0 = 24 ms for fiu_search_batch, 6 ms for
                                the server (requests join between steps) */
    uint32_t candidate_cap;  /* candidate buffer entries per search */
    uint32_t flags;          /* reserved, 0 */
} fiu_batch_config;

void fiu_batch_config_init(fiu_batch_config *cfg);

typedef struct {
    const fiu_dataset *dataset;
    uint16_t order;
    uint16_t reserved;
    int32_t dist_step, dist_offset;   /* Distribution (step, offset) */
    uint32_t outputs;
    int32_t *out_combos;              /* outputs*order ints */
    float *out_mi;                    /* outputs floats */
    uint32_t out_count;               /* filled: results */
    int32_t status;                   /* filled: fiu_status of this search */
} fiu_scan;

typedef struct {
    uint64_t steps;          /* batched steps launched */
    uint64_t units;          /* work units claimed: wavefront units of the
                                popcount kinds, workgroup units (tiles) of
                                the matrix-core kinds */
    uint64_t redo_ranges;    /* prefix ranges scanned again (candidate overflow) */
    double device_ms;        /* device time of the steps (HIP events) */
    double total_ms;         /* wall time of the call */
} fiu_batch_stats;

/* Run n searches together; returns when all are done.  Each fiu_scan gets
 * its own results and status; the call returns the first failing status
 * (FIU_OK when every search succeeded).  `cfg` and `stats` may be NULL. */
fiu_status fiu_search_batch(fiu_scan *scans, size_t n, const fiu_batch_config *cfg,
                            fiu_batch_stats *stats);

/* Scan server: a scheduler thread keeps stepping the searches on the device
 * while new requests are admitted.  A request names a resident dataset, or
 * carries host tables (reference layout, copied at submit) that the server
 * uploads into its memory budget when the request's footprint (tables, pair
 * table, binomials) fits, and frees when the request completes; the fp4
 * image of the matrix-core scans is added only when it fits as well.
 * Requests that can never fit are refused with FIU_ERR_REJECTED. */
typedef struct fiu_server fiu_server;

typedef struct {
    uint32_t struct_size;             /* sizeof(fiu_request) */
    uint16_t order;
    uint16_t reserved;
    const fiu_dataset *dataset;       /* resident dataset, or NULL and: */
    const uint64_t *words;            /* host tables, as fiu_dataset_create */
    uint64_t snps, cases_words, ctrls_words;
    uint32_t cases, ctrls;
    int32_t dist_step, dist_offset;
    uint32_t outputs;
    uint32_t reserved2;
} fiu_request;

typedef struct {
    double submit_ms, admit_ms, start_ms, done_ms;   /* since server creation */
    double latency_ms;       /* done_ms - submit_ms */
    uint64_t combinations;   /* combinations evaluated */
    uint64_t steps;          /* batched steps the request took part in */
    int32_t algorithm;       /* fiu_algorithm used */
    int32_t reserved;
} fiu_request_stats;

typedef struct {
    uint64_t queued;         /* submitted, not yet on the device */
    uint64_t active;         /* on the device */
    uint64_t completed;
    uint64_t memory_budget, memory_used;
    uint64_t steps;
    double device_ms;
} fiu_server_info;

fiu_status fiu_server_create(const fiu_batch_config *cfg, fiu_server **out);
fiu_status fiu_server_submit(fiu_server *sv, const fiu_request *req, uint64_t *id);
/* Wait for request `id` (timeout_ms < 0: no limit; 0: poll) and collect its
 * results (out buffers of outputs*order ints and outputs floats).  Returns
 * FIU_ERR_PENDING on timeout; a collected request is forgotten. */
fiu_status fiu_server_wait(fiu_server *sv, uint64_t id, double timeout_ms,
                           int32_t *out_combos, float *out_mi, uint32_t *out_count,
                           fiu_request_stats *stats);
fiu_status fiu_server_query(fiu_server *sv, fiu_server_info *out);
/* Finishes the requests already on the device; queued ones are dropped. */
fiu_status fiu_server_destroy(fiu_server *sv);


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

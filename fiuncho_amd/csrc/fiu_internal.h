// Host-side internals shared by the C ABI (fiu_api.hip) and the batched scan
// engine / scan server (fiu_engine.hip): error plumbing, device buffers, the
// dataset handle and the per-dataset preparation steps.
#pragma once

#include "../../include/fiuncho_amd.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "fiu_kernels.h"
#include "fiu_loader.h"

namespace fiu {

struct Failure {
    fiu_status status;
    std::string msg;
};

[[noreturn]] inline void fail(fiu_status s, const std::string &m) { throw Failure{s, m}; }

inline void check(hipError_t e, const char *what)
{
    if (e != hipSuccess) fail(FIU_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// Thread-local message slot behind fiu_last_error(), and the device selected
// by fiu_set_device() on this thread.
std::string &error_slot();
int &current_device();

template <class F> fiu_status guard(F &&f)
{
    try {
        f();
        error_slot().clear();
        return FIU_OK;
    } catch (const Failure &e) {
        error_slot() = e.msg;
        return e.status;
    } catch (const LoadError &e) {
        error_slot() = e.what();
        return FIU_ERR_IO;
    } catch (const std::bad_alloc &) {
        error_slot() = "host allocation failed";
        return FIU_ERR_NOMEM;
    } catch (const std::exception &e) {
        error_slot() = e.what();
        return FIU_ERR_INVALID;
    }
}

// First-fit sub-allocator over one device allocation (or one pinned host
// allocation): the scan server takes its whole memory budget once and carves
// tables and buffers out of it, so that admitting and retiring scans never
// calls hipMalloc / hipFree (which can wait for the device).  Thread-safe.
class Arena {
  public:
    static constexpr size_t kAlign = 256;
    Arena(size_t bytes, int device, bool pinned_host = false);
    ~Arena();
    Arena(const Arena &) = delete;
    Arena &operator=(const Arena &) = delete;
    void *alloc(size_t bytes);   // nullptr when no free range fits
    void release(void *p);
    size_t capacity() const { return size_; }
    size_t used() const;
    size_t largest_free() const;

  private:
    char *base_ = nullptr;
    bool pinned_ = false;
    size_t size_ = 0;
    size_t used_ = 0;
    std::map<size_t, size_t> free_;   // offset -> length
    std::map<size_t, size_t> taken_;  // offset -> length
    mutable std::mutex mu_;
};

// Device buffer with RAII, from hipMalloc or from an Arena.
template <class T> struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    Arena *arena = nullptr;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    void release()
    {
        if (p) {
            if (arena) arena->release(p);
            else (void)hipFree(p);
        }
        p = nullptr;
        n = 0;
    }
    // Hand the allocation over to the caller (who frees it later, through
    // `arena` when set) and forget it.
    T *detach()
    {
        T *q = p;
        p = nullptr;
        n = 0;
        return q;
    }
    void alloc(size_t count)
    {
        if (count <= n && p) return;
        release();
        const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
        if (arena) {
            p = (T *)arena->alloc(bytes);
            if (!p) fail(FIU_ERR_NOMEM, "scan server memory budget exhausted (" + std::to_string(bytes) + " bytes)");
        } else if (hipMalloc(&p, bytes) != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            fail(FIU_ERR_NOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed");
        }
        n = count;
    }
};

// Every host-issued copy names the allocation it touches: a range that does
// not lie inside it is a bug in the caller, reported as FIU_ERR_DEVICE
// instead of being handed to the runtime.
inline void check_range(const void *p, size_t bytes, const void *base, size_t cap, const char *what)
{
    const uintptr_t a = (uintptr_t)p, b = (uintptr_t)base;
    if (!p || !base || a < b || a - b > cap || bytes > cap - (a - b))
        fail(FIU_ERR_DEVICE, std::string("internal: ") + what + " copy outside its allocation");
}

uint64_t binom_sat(uint64_t n, uint64_t k);
// Lexicographic unranking of an m-combination of [0, n) (ascending indices).
void unrank_host(uint64_t r, uint64_t n, int m, int32_t *out);

// Exact number of combinations whose m-prefix has lexicographic rank
// = offset (mod step).
uint64_t count_combinations(uint64_t n, int m, uint64_t step, uint64_t offset);

// Tuning table (fiu_set_tuning): the entry of the class holding (order,
// snps, samples), if any.
bool tuned(int order, uint64_t snps, uint64_t samples, fiu_tune_entry &out);
// `variant` is one of `algo`'s kernel variants at `order` (0 = default).
bool valid_variant(int algo, int order, int variant);

// H(Y), 1/N and the screening margin of a dataset's class sizes.  `complete`:
// every sample has a genotype in {0, 1, 2} at every SNP, so that a table's
// cells sum to the class sizes and the MI upper bound applies.
void mi_constants(uint32_t cases, uint32_t ctrls, bool complete, ScanParams &P);

}  // namespace fiu

struct fiu_dataset {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = true;
    fiu::Arena *arena = nullptr;   // set: every device buffer comes from it
    size_t snps = 0, cw = 0, tw = 0;
    uint32_t cases = 0, ctrls = 0;
    std::vector<uint64_t> host_words;
    fiu::DBuf<uint4> bits;
    fiu::DBuf<uint32_t> singles;
    fiu::DBuf<uint32_t> flags;
    fiu::DevData dev{};
    // search scratch (grown on demand, kept for later calls)
    fiu::DBuf<uint64_t> binom;
    int binom_m = -1;
    fiu::DBuf<fiu::Cand> cand, top;
    fiu::DBuf<fiu::TopState> state;
    fiu::DBuf<uint32_t> redo;
    fiu::DBuf<uint32_t> pair;        // marginal pair table (order-3 marginal scans)
    fiu::DBuf<uint32_t> hist;        // screening-MI histogram of the threshold sampling pass
    fiu::DBuf<uint4> fp4;            // fp4 image of genotype rows 0/1 (matrix-core scans)
    fiu::DBuf<uint32_t> big;         // orders 10-12: count slots (big_cells_bytes)
    fiu::DBuf<fiu::Cand> mwork;      // outputs above kMaxOutputs: the merge's sort buffer
    bool fp4_ready = false;
    bool pair_ready = false;
    std::vector<hipEvent_t> events;
    void *readback = nullptr;        // pinned: search state + up to kMaxOutputs results
    std::mutex prep_mu;              // binom / pair table preparation
    // device buffers replaced while work queued earlier may still read them
    // (a binomial table grown for a higher order): freed with the dataset
    std::vector<std::pair<void *, fiu::Arena *>> retired;

    // every sample has a genotype in {0, 1, 2} at every SNP: the cells of a
    // table sum to the class sizes
    bool complete() const { return dev.uniform && dev.total[0] == cases && dev.total[1] == ctrls; }

    void use_arena(fiu::Arena *a)
    {
        arena = a;
        bits.arena = singles.arena = flags.arena = binom.arena = redo.arena = pair.arena = hist.arena = a;
        fp4.arena = big.arena = mwork.arena = a;
        cand.arena = top.arena = a;
        state.arena = a;
    }
    ~fiu_dataset();
};

namespace fiu {

// Copy host-layout words to the device image of a dataset (transpose to the
// SNP-minor layout, row popcounts, coverage flags).  `stream` null: the
// dataset gets a stream of its own.  Memory from `arena` when given.
fiu_dataset *create_dataset(const uint64_t *words, size_t snps, size_t cw, size_t tw, uint32_t cases,
                            uint32_t ctrls, Arena *arena = nullptr, hipStream_t stream = nullptr,
                            bool keep_host_words = true);

// Pack a SNP-major genotype matrix on the device into the reference layout
// (as RawFile::fill does; `reader_flush`: with the readers' re-flushed words
// zeroed) and build the device image from it.
fiu_dataset *dataset_from_genotypes(const int8_t *geno, size_t snps, size_t samples, const uint8_t *is_ctrl,
                                    bool reader_flush, Arena *arena = nullptr, hipStream_t stream = nullptr);

// Device bytes a dataset of this shape takes (bits, row counts, flags) and
// the transient staging its upload needs.
size_t dataset_bytes(size_t snps, size_t cw, size_t tw);
size_t staging_bytes(size_t snps, size_t cw, size_t tw);

// Staged upload (scan server): the host does every part of the preparation
// that needs no device (row counts, coverage check, binomials) and copies the
// tables into pinned memory; enqueue_upload() later puts the copies, the
// transposes and the pair table on one stream with no host synchronisation,
// so that preparing a dataset never waits for the work already queued.
struct StagedUpload {
    char *host = nullptr;            // pinned: [words][row counts][binomials]
    size_t host_bytes = 0;
    size_t n_words = 0, off_singles = 0, off_binom = 0, binom_n = 0;
    uint64_t *staging_d = nullptr;   // device copy of the words (arena)
    bool pair = false;               // build the order-3 pair table
    bool fp4 = false;                // build the fp4 image (matrix-core scans)
};

// Host bytes (pinned) a staged upload of this shape needs for `order`.
size_t staged_host_bytes(size_t snps, size_t cw, size_t tw, int order);

// Allocates the dataset's device buffers (and the device staging) from
// `arena` and its pinned staging from `pinned`; returns null when either has
// no room.  Throws on invalid shapes.  `want_fp4`: also the fp4 image when the
// dataset qualifies for the matrix-core scans (mfma_dataset; order 3 also
// needs the pair table) and it fits.
fiu_dataset *stage_dataset(const uint64_t *words, size_t snps, size_t cw, size_t tw, uint32_t cases,
                           uint32_t ctrls, int order, bool want_fp4, Arena *arena, Arena *pinned, StagedUpload &up);
// Copies, transposes and builds the pair table on `s`.  Both stagings --
// the pinned host block up.host and the device block up.staging_d -- are
// read by that queued work: the caller releases them only once it has
// completed (the scan engine releases them after the next step on `s`).
void enqueue_upload(fiu_dataset *d, StagedUpload &up, hipStream_t s);

// binom[j*(n+1)+x] = C(x, j) for j <= m, on the device; returns the table.
// A table grown for a higher order replaces the dataset's current one, which
// stays allocated (work queued earlier may read it) until the dataset is
// destroyed, so a returned pointer stays valid for the dataset's lifetime.
const uint64_t *ensure_binom(fiu_dataset *d, int m, hipStream_t s);
// Pair counts f(x, y), x, y in {0, 1}, of every SNP pair (32 bytes per
// pair).  Returns false when the table does not fit in free device memory
// (hipMalloc datasets) or in the arena.
bool ensure_pair_table(fiu_dataset *d, hipStream_t s);
inline size_t pair_table_bytes(size_t snps) { return snps < 2 ? 0 : snps * (snps - 1) / 2 * 32; }
// fp4 image of genotype rows 0 and 1 (4 bits per sample and row); false when
// it does not fit.
bool ensure_fp4_image(fiu_dataset *d, hipStream_t s);
// The dataset admits the matrix-core scans: every SNP covers the same samples
// (the counts of the other cells follow by inclusion-exclusion) and both
// classes are below 2^24 samples (FP32 counts are exact).
inline bool mfma_dataset(const fiu_dataset *d)
{
    return d->dev.uniform && d->cases < (1u << 24) && d->ctrls < (1u << 24);
}

// Top-N entries -> ascending SNP indices (order per entry, unranked from the
// key among the k-combinations of n SNPs) and float MI.
void unpack_results(const Cand *top, uint32_t count, int order, uint64_t n, int32_t *out_combos, float *out_mi);

}  // namespace fiu

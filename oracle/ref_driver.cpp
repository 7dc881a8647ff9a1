// C entry points over the reference's own search path, compiled from the
// reference sources where they lie (see oracle/Makefile) into oracle/_ref/.
//
// TEST INFRASTRUCTURE ONLY: used to write golden vectors (tests/golden/) and,
// when present, as the timed CPU baseline.  It includes the reference headers
// unchanged and calls Dataset<uint64_t>::read (Dataset.h:48-94),
// GenotypeTable::combine / combine_and_popcnt (GenotypeTable.h:123-140),
// MutualInformation<float>::compute (MutualInformation.h:83-84) and
// ThreadedSearch::run (ThreadedSearch.h:207-246).
#include <fiuncho/ContingencyTable.h>
#include <fiuncho/Dataset.h>
#include <fiuncho/Distribution.h>
#include <fiuncho/GenotypeTable.h>
#include <fiuncho/ThreadedSearch.h>
#include <fiuncho/algorithms/MutualInformation.h>
#include <fiuncho/utils/Result.h>

#include <chrono>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

namespace {
struct Handle {
    Dataset<uint64_t> d;
};
thread_local std::string g_err;
} // namespace

extern "C" {

const char *ref_last_error() { return g_err.c_str(); }

// Load one .raw file or a .tped/.tfam pair (any order).
void *ref_load(const char *f1, const char *f2)
{
    try {
        std::vector<std::string> in{f1};
        if (f2 && *f2) in.emplace_back(f2);
        return new Handle{Dataset<uint64_t>::read(in)};
    } catch (const std::exception &e) {
        g_err = e.what();
        return nullptr;
    }
}

void ref_free(void *h) { delete static_cast<Handle *>(h); }

void ref_info(void *h, size_t *snps, size_t *cases, size_t *ctrls,
              size_t *cases_words, size_t *ctrls_words)
{
    auto &d = static_cast<Handle *>(h)->d;
    *snps = d.snps;
    *cases = d.cases;
    *ctrls = d.ctrls;
    *cases_words = d[0].cases_words;
    *ctrls_words = d[0].ctrls_words;
}

// Copy every SNP table (3 case rows then 3 control rows per SNP).
void ref_words(void *h, uint64_t *out)
{
    auto &d = static_cast<Handle *>(h)->d;
    const size_t cw = d[0].cases_words, tw = d[0].ctrls_words;
    for (size_t s = 0; s < d.snps; ++s) {
        std::memcpy(out, d[s].cases, 3 * cw * sizeof(uint64_t));
        out += 3 * cw;
        std::memcpy(out, d[s].ctrls, 3 * tw * sizeof(uint64_t));
        out += 3 * tw;
    }
}

// Counts and float MI of one ascending combination, through the same prefix
// chain as ThreadedSearch.h:160-178.  ct_out receives ct.size cases then
// ct.size controls; returns ct.size.
size_t ref_combination(void *h, const int *combo, int k, uint32_t *ct_out,
                       float *mi_out)
{
    auto &d = static_cast<Handle *>(h)->d;
    std::vector<GenotypeTable<uint64_t>> gts;
    gts.reserve(k);
    for (int o = 2; o < k; ++o)
        gts.emplace_back(o, d[0].cases_words, d[0].ctrls_words);
    for (int i = 0; i < k - 2; ++i)
        GenotypeTable<uint64_t>::combine(i > 0 ? gts[i - 1] : d[combo[0]],
                                         d[combo[i + 1]], gts[i]);
    ContingencyTable<uint32_t> ct(k);
    GenotypeTable<uint64_t>::combine_and_popcnt(
        k > 2 ? gts.back() : d[combo[0]], d[combo[k - 1]], ct);
    MutualInformation<float> mi(d.cases, d.ctrls);
    *mi_out = mi.compute(ct);
    std::memcpy(ct_out, ct.cases, ct.size * sizeof(uint32_t));
    std::memcpy(ct_out + ct.size, ct.ctrls, ct.size * sizeof(uint32_t));
    return ct.size;
}

// Output of GenotypeTable::combine for t1 = SNP a, t2 = SNP b (9 rows per
// class): out receives 9*cw case words then 9*tw control words.
void ref_combine_pair(void *h, int a, int b, uint64_t *out)
{
    auto &d = static_cast<Handle *>(h)->d;
    GenotypeTable<uint64_t> r(2, d[0].cases_words, d[0].ctrls_words);
    GenotypeTable<uint64_t>::combine(d[a], d[b], r);
    std::memcpy(out, r.cases, 9 * r.cases_words * sizeof(uint64_t));
    std::memcpy(out + 9 * r.cases_words, r.ctrls,
                9 * r.ctrls_words * sizeof(uint64_t));
}

// MutualInformation<float>::compute on a caller-given table of 3^k cells per
// class (padding cells are zeroed here, as ContingencyTable holds them).
float ref_mi(const uint32_t *cases_ct, const uint32_t *ctrls_ct, int k,
             unsigned cases, unsigned ctrls)
{
    ContingencyTable<uint32_t> ct(k);
    size_t cells = 1;
    for (int i = 0; i < k; ++i) cells *= 3;
    for (size_t i = 0; i < ct.size; ++i) {
        ct.cases[i] = i < cells ? cases_ct[i] : 0;
        ct.ctrls[i] = i < cells ? ctrls_ct[i] : 0;
    }
    MutualInformation<float> mi(cases, ctrls);
    return mi.compute(ct);
}

// ThreadedSearch(threads).run(dataset, order, Distribution(snps, order-1,
// step, offset), outputs).  Returns the result count; seconds gets the wall
// time of run() alone.
long ref_search(void *h, int order, int step, int offset, unsigned outputs,
                unsigned threads, int *combos, float *mi, double *seconds)
{
    auto &d = static_cast<Handle *>(h)->d;
    try {
        Distribution<int> dist((int)d.snps, order - 1, step, offset);
        ThreadedSearch search(threads);
        auto t0 = std::chrono::steady_clock::now();
        auto res = search.run(d, (unsigned short)order, dist, outputs);
        auto t1 = std::chrono::steady_clock::now();
        if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
        for (size_t i = 0; i < res.size(); ++i) {
            for (int q = 0; q < order; ++q)
                combos[i * order + q] = res[i].combination[q];
            mi[i] = res[i].val;
        }
        return (long)res.size();
    } catch (const std::exception &e) {
        g_err = e.what();
        return -1;
    }
}

// Result<int, float>::str (Result.h:37-45) and Result::serialize
// (Result.h:47-55) of one result: copies at most cap bytes into buf (str adds
// no terminator) and returns the full size.
static size_t copy_out(const std::string &s, char *buf, size_t cap)
{
    if (buf) std::memcpy(buf, s.data(), s.size() < cap ? s.size() : cap);
    return s.size();
}

size_t ref_result_str(const int *combo, int k, float val, char *buf, size_t cap)
{
    Result<int, float> r;
    r.combination.assign(combo, combo + k);
    r.val = val;
    return copy_out(r.str(), buf, cap);
}

size_t ref_result_serialize(const int *combo, int k, float val, char *buf, size_t cap)
{
    Result<int, float> r;
    r.combination.assign(combo, combo + k);
    r.val = val;
    std::ostringstream os;
    Result<int, float>::serialize(os, r);
    return copy_out(os.str(), buf, cap);
}

} // extern "C"

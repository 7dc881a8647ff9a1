/*
 * GpuSearch -- the reference's Search interface (include/fiuncho/Search.h:
 * 19-46) implemented over the C ABI of libfiuncho_amd.so (fiuncho_amd.h).
 *
 * A drop-in for ThreadedSearch (ThreadedSearch.h:28-247) wherever the
 * reference instantiates a Search: MPIEngine::run<GpuSearch>(inputs, order,
 * outputs, device) (MPIEngine.h:135-137), or directly, as
 * tests/threadedsearch.cpp:19-22 does with ThreadedSearch.  Compile with the
 * reference's include directory on the include path and link
 * libfiuncho_amd.so.
 *
 * Semantics (Search.h:31-43):
 *   - the dataset is Dataset<uint64_t>::read's: one contiguous block of
 *     GenotypeTable rows (GenotypeTable.h:77-101), uploaded as it is;
 *   - the search evaluates exactly the combinations whose (k-1)-prefix belongs
 *     to `distribution` (Distribution.h:147-176), the set ThreadedSearch(1)
 *     evaluates, so MPI rank splits and layer() keep their meaning;
 *   - results: at most `outputs`, MI descending (ties by combination
 *     ascending), SNP indices ascending within a result, MI as float
 *     (ThreadedSearch.h:241-244);
 *   - errors: std::runtime_error with the library's message, as the
 *     reference's I/O and validation errors are (app/main.cpp:163-166).
 */
#ifndef FIUNCHO_AMD_SEARCH_H
#define FIUNCHO_AMD_SEARCH_H

#include <fiuncho/Dataset.h>
#include <fiuncho/Distribution.h>
#include <fiuncho/Search.h>
#include <fiuncho/utils/Result.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "fiuncho_amd.h"

class GpuSearch : public Search
{
  public:
    /* device: the HIP device to search on; algorithm: FIU_ALGO_AUTO (the
     * fastest exact path the dataset allows) or a fixed fiu_algorithm. */
    explicit GpuSearch(int device = 0, int algorithm = FIU_ALGO_AUTO)
    {
        fiu_config_init(&cfg_);
        cfg_.algorithm = algorithm;
        if (fiu_set_device(device) != FIU_OK) throw std::runtime_error(fiu_last_error());
    }

    ~GpuSearch() override { unbind(); }
    GpuSearch(const GpuSearch &) = delete;
    GpuSearch &operator=(const GpuSearch &) = delete;

    /* Keep `dataset` resident on the device for the following run() calls on
     * that same object (the caller keeps it alive and unchanged meanwhile);
     * run() otherwise uploads the tables for each call. */
    void bind(const Dataset<uint64_t> &dataset)
    {
        unbind();
        ds_ = upload(dataset);
        bound_ = &dataset;
    }

    void unbind()
    {
        if (ds_) fiu_dataset_destroy(ds_);
        ds_ = nullptr;
        bound_ = nullptr;
    }

    std::vector<Result<int, float>> run(const Dataset<uint64_t> &dataset, const unsigned short order,
                                        const Distribution<int> &distribution,
                                        const unsigned int outputs) override
    {
        fiu_dataset *ds = &dataset == bound_ ? ds_ : upload(dataset);
        std::vector<int32_t> combos((size_t)outputs * order);
        std::vector<float> mi(outputs);
        uint32_t count = 0;
        const fiu_status s = fiu_search(ds, order, distribution.step, distribution.offset, outputs, &cfg_,
                                        combos.data(), mi.data(), &count, &stats_);
        const std::string err = s == FIU_OK ? std::string() : std::string(fiu_last_error());
        if (ds != ds_) fiu_dataset_destroy(ds);
        if (s != FIU_OK) throw std::runtime_error(err);
        std::vector<Result<int, float>> out(count);
        for (uint32_t i = 0; i < count; ++i) {
            out[i].combination.assign(combos.begin() + (size_t)i * order, combos.begin() + (size_t)(i + 1) * order);
            out[i].val = mi[i];
        }
        return out;
    }

    /* Counters of the last run() (combinations, launches, device time, the
     * algorithm used). */
    const fiu_search_stats &last_stats() const { return stats_; }

  private:
    static fiu_dataset *upload(const Dataset<uint64_t> &dataset)
    {
        if (dataset.snps == 0) throw std::runtime_error("empty dataset");
        // one contiguous allocation: SNP s at s * 3 * (cw + tw) words from
        // SNP 0's case rows (GenotypeTable::make_array)
        const size_t cw = dataset[0].cases_words, tw = dataset[0].ctrls_words;
        fiu_dataset *ds = nullptr;
        if (fiu_dataset_create(dataset[0].cases, dataset.snps, cw, tw, (uint32_t)dataset.cases,
                               (uint32_t)dataset.ctrls, &ds) != FIU_OK)
            throw std::runtime_error(fiu_last_error());
        return ds;
    }

    fiu_config cfg_;
    fiu_search_stats stats_{};
    fiu_dataset *ds_ = nullptr;
    const Dataset<uint64_t> *bound_ = nullptr;
};

#endif

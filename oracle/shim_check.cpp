// GpuSearch (include/fiuncho_amd_search.h) used as the reference uses a
// Search: the dataset read by Dataset<uint64_t>::read (Dataset.h:48-94), then
// run() as tests/threadedsearch.cpp:17-31 calls ThreadedSearch::run, next to
// ThreadedSearch(1) on the same calls.
//
// TEST INFRASTRUCTURE ONLY: compiled from the reference's headers and sources
// where they lie (oracle/Makefile, target `shim`) into oracle/_ref/; run by
// tests/test_gpu_boundary.py on a GPU.
//
// Usage: shim_check FILE [FILE]   (a .raw, or a .tped and its .tfam)
// Prints "order K step S offset O: N results" per case, the Result::str
// lines of the order-3 results, and "shim ok"; exit status 1 when a GpuSearch
// list differs from ThreadedSearch's beyond ties (entries in only one list
// within 4e-6 of the last value; common entries within 1e-5 relative + 2e-6).
#include <fiuncho/Dataset.h>
#include <fiuncho/Distribution.h>
#include <fiuncho/ThreadedSearch.h>

#include <cmath>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "../include/fiuncho_amd_search.h"

static bool same_up_to_ties(const std::vector<Result<int, float>> &a, const std::vector<Result<int, float>> &b)
{
    if (a.size() != b.size()) return false;
    if (a.empty()) return true;
    std::map<std::vector<int>, float> ma, mb;
    for (auto &r : a) ma[r.combination] = r.val;
    for (auto &r : b) mb[r.combination] = r.val;
    float last = b.back().val;
    for (auto &r : b) last = std::min(last, r.val);
    for (auto &kv : ma) {
        auto it = mb.find(kv.first);
        if (it == mb.end()) {
            if (std::fabs(kv.second - last) > 4e-6f) return false;
        } else if (std::fabs(kv.second - it->second) > 1e-5f * std::fabs(it->second) + 2e-6f) {
            return false;
        }
    }
    for (auto &kv : mb)
        if (!ma.count(kv.first) && std::fabs(kv.second - last) > 4e-6f) return false;
    // order: MI descending, then combination ascending
    for (size_t i = 1; i < a.size(); ++i)
        if (a[i].val > a[i - 1].val || (a[i].val == a[i - 1].val && !(a[i - 1].combination < a[i].combination)))
            return false;
    return true;
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s FILE [FILE]\n", argv[0]);
        return 2;
    }
    std::vector<std::string> files(argv + 1, argv + argc);
    try {
        const Dataset<uint64_t> dataset = Dataset<uint64_t>::read(files);
        const int n = (int)dataset.snps;
        GpuSearch gpu(0);
        bool ok = true;
        for (int order = 2; order <= 4; ++order) {
            for (auto so : {std::make_pair(1, 0), std::make_pair(3, 1)}) {
                const Distribution<int> dist(n, order - 1, so.first, so.second);
                if (so.first == 3 && order == 3) gpu.bind(dataset);   // resident from here on
                auto g = gpu.run(dataset, (unsigned short)order, dist, 90);
                auto r = ThreadedSearch(1).run(dataset, (unsigned short)order, dist, 90);
                const bool same = same_up_to_ties(g, r);
                std::printf("order %d step %d offset %d: %zu results%s\n", order, so.first, so.second, g.size(),
                            same ? "" : "  MISMATCH");
                ok = ok && same && !g.empty();
                if (order == 3 && so.first == 1)
                    for (auto &res : g) std::printf("%s\n", res.str().c_str());
            }
        }
        gpu.unbind();
        std::printf(ok ? "shim ok\n" : "shim FAILED\n");
        return ok ? 0 : 1;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}

// A C++ program that uses libfiuncho_amd.so through its C ABI only
// (include/fiuncho_amd.h), as a non-Python caller does.
//
//   abi_caller host   host-only entry points: version, table sizes, the
//                     synthetic generator, the result formats
//   abi_caller gpu    also a search on GPU 0: prints the top-10 pairs and
//                     top-5 triples of a seeded synthetic dataset as
//                     Result::str lines
// Exit status 0 when every check passes.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "fiuncho_amd.h"

#define CHECK(cond)                                                              \
    do {                                                                         \
        if (!(cond)) {                                                           \
            std::fprintf(stderr, "check failed: %s (%s)\n", #cond, fiu_last_error()); \
            return 1;                                                            \
        }                                                                        \
    } while (0)

static int host_checks()
{
    CHECK(std::strstr(fiu_version(), "gfx950") != nullptr);
    CHECK(fiu_ct_size(2) == 16 && fiu_ct_size(3) == 32 && fiu_ct_size(4) == 96);
    fiu_synth_spec spec;
    fiu_synth_spec_init(&spec);
    spec.snps = 20;
    spec.cases = 100;
    spec.ctrls = 90;
    size_t cw = 0, tw = 0;
    CHECK(fiu_synth_generate(&spec, nullptr, nullptr, &cw, &tw) == FIU_OK && cw == 8 && tw == 8);
    std::vector<uint64_t> words(spec.snps * 3 * (cw + tw));
    CHECK(fiu_synth_generate(&spec, words.data(), nullptr, &cw, &tw) == FIU_OK);
    const int32_t combo[3] = {4, 8, 9};
    char line[64];
    const size_t n = fiu_result_str(combo, 3, 0.137897015f, line, sizeof line);
    CHECK(n == std::strlen("4 8 9 0.137897") && std::strcmp(line, "4 8 9 0.137897") == 0);
    unsigned char rec[64];
    const size_t rb = fiu_result_serialize(combo, 3, 0.5f, rec, sizeof rec);
    CHECK(rb == sizeof(size_t) + 3 * 4 + 4);
    int32_t back[8];
    uint16_t order = 0;
    float mi = 0;
    size_t used = 0;
    CHECK(fiu_result_deserialize(rec, rb, back, 8, &order, &mi, &used) == FIU_OK);
    CHECK(order == 3 && used == rb && mi == 0.5f && back[0] == 4 && back[1] == 8 && back[2] == 9);
    CHECK(fiu_result_deserialize(rec, rb - 1, back, 8, &order, &mi, &used) == FIU_ERR_INVALID);
    return 0;
}

static int gpu_search()
{
    CHECK(fiu_set_device(0) == FIU_OK);
    fiu_synth_spec spec;
    fiu_synth_spec_init(&spec);
    spec.snps = 300;
    spec.cases = 1000;
    spec.ctrls = 900;
    spec.seed = 5;
    spec.planted_order = 2;
    spec.planted[0] = 7;
    spec.planted[1] = 250;
    fiu_dataset *ds = nullptr;
    CHECK(fiu_dataset_synth(&spec, &ds) == FIU_OK);
    for (int order = 2; order <= 3; ++order) {
        const uint32_t outputs = order == 2 ? 10 : 5;
        std::vector<int32_t> combos(outputs * order);
        std::vector<float> mi(outputs);
        uint32_t count = 0;
        fiu_search_stats st;
        CHECK(fiu_search(ds, (uint16_t)order, 1, 0, outputs, nullptr, combos.data(), mi.data(), &count, &st) ==
              FIU_OK);
        CHECK(count == outputs);
        for (uint32_t i = 0; i < count; ++i) {
            char line[128];
            fiu_result_str(&combos[(size_t)i * order], (uint16_t)order, mi[i], line, sizeof line);
            std::printf("%s\n", line);
        }
        if (order == 2) CHECK(combos[0] == 7 && combos[1] == 250);
    }
    CHECK(fiu_dataset_destroy(ds) == FIU_OK);
    return 0;
}

int main(int argc, char **argv)
{
    const std::string mode = argc > 1 ? argv[1] : "host";
    if (host_checks() != 0) return 1;
    if (mode == "gpu" && gpu_search() != 0) return 1;
    std::printf("abi_caller ok\n");
    return 0;
}

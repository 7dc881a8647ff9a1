// Seeded synthetic genotype data (see fiu_synth.cpp for the random stream).
#pragma once

#include <stddef.h>
#include <stdint.h>

namespace fiu {

struct SynthSpec {
    uint64_t snps = 0;
    uint32_t cases = 0, ctrls = 0;
    double maf_lo = 0.05, maf_hi = 0.5;
    double low_frac = 0.0, low_lo = 0.01, low_hi = 0.05;
    uint64_t seed = 1;
    int planted_order = 0;
    int64_t planted[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint64_t pen_seed = 2;
    double pen_lo = 0.3, pen_hi = 0.7;
};

double synth_maf(const SynthSpec &s, uint64_t snp);
size_t synth_words(uint64_t n);
// words: snps*3*(cw+tw) (reference layout) or null; geno: snps x samples or null.
void synth_generate(const SynthSpec &s, uint64_t *words, int8_t *geno, int threads);

}  // namespace fiu

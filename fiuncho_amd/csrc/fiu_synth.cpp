// Seeded synthetic case/control genotype data (Hardy-Weinberg genotypes,
// optional planted epistatic combination), written straight into the
// reference's bit-packed layout.
//
// Random stream: every draw is u(seed, stream, counter) = mix(seed*K1 +
// stream*K2 + counter) / 2^53 with mix the SplitMix64 finaliser (Steele,
// Lea & Flood, "Fast splittable pseudorandom number generators", 2014), so
// any SNP or sample can be regenerated independently and in parallel.
//   stream s < snps, counter 0      MAF draw of SNP s
//   stream s, counter 1             MAF bucket draw (skewed spectra)
//   stream s, counter 2 + j         genotype of sample j
//   stream 2^40 + i                 penetrance cell i (seeded by pen seed)
//   stream 2^41, counter j          joint genotype of the planted SNPs
// Samples 0..cases-1 are cases, the rest controls (labels fixed by index).
// For a planted combination the joint genotype of its SNPs in a case is drawn
// from P(g) pen(g) / Z and in a control from P(g) (1 - pen(g)) / Z', with
// P(g) the product of the SNPs' HWE probabilities.
#include "fiu_synth.h"

#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

namespace fiu {
namespace {

constexpr uint64_t K1 = 0x9E3779B97F4A7C15ull, K2 = 0xD1B54A32D192ED03ull;

inline uint64_t mix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline double uni(uint64_t seed, uint64_t stream, uint64_t ctr)
{
    return (double)(mix(seed * K1 + stream * K2 + ctr) >> 11) * (1.0 / 9007199254740992.0);
}

inline int hwe_draw(double p, double u)
{
    const double q = 1.0 - p;
    return u < q * q ? 0 : (u < q * q + 2.0 * p * q ? 1 : 2);
}

}  // namespace

double synth_maf(const SynthSpec &s, uint64_t snp)
{
    if (s.low_frac > 0.0 && uni(s.seed, snp, 1) < s.low_frac)
        return s.low_lo + (s.low_hi - s.low_lo) * uni(s.seed, snp, 0);
    return s.maf_lo + (s.maf_hi - s.maf_lo) * uni(s.seed, snp, 0);
}

size_t synth_words(uint64_t n) { return (n + 511) / 512 * 8; }

void synth_generate(const SynthSpec &s, uint64_t *words, int8_t *geno, int threads)
{
    const uint64_t N = (uint64_t)s.cases + s.ctrls;
    const size_t cw = synth_words(s.cases), tw = synth_words(s.ctrls);
    const size_t per = 3 * (cw + tw);
    // planted joint genotypes, one per sample
    std::vector<uint8_t> joint;
    const int K = s.planted_order;
    if (K > 0) {
        int cells = 1;
        for (int i = 0; i < K; ++i) cells *= 3;
        std::vector<double> pen(cells), pc(cells), pt(cells);
        for (int i = 0; i < cells; ++i)
            pen[i] = s.pen_lo + (s.pen_hi - s.pen_lo) * uni(s.pen_seed, (1ull << 40) + i, 0);
        double zc = 0, zt = 0;
        for (int cell = 0; cell < cells; ++cell) {
            double p = 1.0;
            int r = cell;
            for (int q = K - 1; q >= 0; --q) {
                const int g = r % 3;
                r /= 3;
                const double m = synth_maf(s, (uint64_t)s.planted[q]), o = 1.0 - m;
                p *= g == 0 ? o * o : (g == 1 ? 2.0 * m * o : m * m);
            }
            pc[cell] = p * pen[cell];
            pt[cell] = p * (1.0 - pen[cell]);
            zc += pc[cell];
            zt += pt[cell];
        }
        for (int cell = 1; cell < cells; ++cell) pc[cell] += pc[cell - 1], pt[cell] += pt[cell - 1];
        joint.resize(N);
        for (uint64_t j = 0; j < N; ++j) {
            const bool is_case = j < s.cases;
            const double u = uni(s.seed, 1ull << 41, j) * (is_case ? zc : zt);
            const std::vector<double> &cdf = is_case ? pc : pt;
            int cell = (int)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
            joint[j] = (uint8_t)std::min(cell, cells - 1);
        }
    }
    auto work = [&](uint64_t s0, uint64_t s1) {
        std::vector<int8_t> g(N);
        for (uint64_t snp = s0; snp < s1; ++snp) {
            int slot = -1;
            for (int q = 0; q < K; ++q)
                if ((uint64_t)s.planted[q] == snp) slot = q;
            if (slot >= 0) {
                int div = 1;
                for (int q = K - 1; q > slot; --q) div *= 3;
                for (uint64_t j = 0; j < N; ++j) g[j] = (int8_t)((joint[j] / div) % 3);
            } else {
                const double p = synth_maf(s, snp);
                for (uint64_t j = 0; j < N; ++j) g[j] = (int8_t)hwe_draw(p, uni(s.seed, snp, 2 + j));
            }
            if (geno) std::copy(g.begin(), g.end(), geno + snp * N);
            if (words) {
                uint64_t *base = words + snp * per;
                std::fill(base, base + per, 0ull);
                for (int cls = 0; cls < 2; ++cls) {
                    const uint64_t j0 = cls ? s.cases : 0, j1 = cls ? N : s.cases;
                    const size_t W = cls ? tw : cw;
                    uint64_t *rows = base + (cls ? 3 * cw : 0);
                    uint64_t acc[3] = {0, 0, 0};
                    uint64_t n = 0;
                    for (uint64_t j = j0; j < j1; ++j) {
                        for (int z = 0; z < 3; ++z) acc[z] = (acc[z] << 1) | (uint64_t)(g[j] == z);
                        if (++n % 64 == 0)
                            for (int z = 0; z < 3; ++z) rows[z * W + n / 64 - 1] = acc[z], acc[z] = 0;
                    }
                    if (n % 64)
                        for (int z = 0; z < 3; ++z) rows[z * W + n / 64] = acc[z];
                }
            }
        }
    };
    const int T = std::max(1, std::min<int>(threads, (int)std::max<uint64_t>(1, s.snps / 16)));
    std::vector<std::thread> pool;
    for (int t = 0; t < T; ++t) {
        const uint64_t a = s.snps * t / T, b = s.snps * (t + 1) / T;
        pool.emplace_back(work, a, b);
    }
    for (auto &th : pool) th.join();
}

}  // namespace fiu

// Host-side genotype file readers (see fiu_loader.cpp).
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace fiu {

struct LoadError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// Bit tables in the reference host layout: SNP s at s*3*(cw+tw) words, three
// case rows of cw words then three control rows of tw words.
struct PackedDataset {
    size_t snps = 0, cases = 0, ctrls = 0, cases_words = 0, ctrls_words = 0;
    std::vector<uint64_t> words;
};

// A parsed genotype file: SNP-major genotypes (snps x samples; 0/1/2, any
// other value = uncalled) and the class of every sample in file order.
struct Genotypes {
    size_t snps = 0, samples = 0;
    std::vector<int8_t> geno;
    std::vector<uint8_t> is_ctrl;
};

// Pack a SNP-major genotype matrix (values 0/1/2, anything else sets no bit).
void pack_genotypes(const std::vector<int8_t> &geno, const std::vector<uint8_t> &is_ctrl, size_t snps,
                    size_t samples, PackedDataset &out);
// (class, word) pairs the reference's file readers re-flush to zero.
std::vector<std::pair<int, size_t>> flush_quirk_words(const uint8_t *is_ctrl, size_t samples);
// 64-bit words per row of a class of n samples (rows padded to 512 bits).
inline size_t packed_words(size_t n) { return (n + 511) / 512 * 8; }
void parse_raw(const std::string &path, Genotypes &out);
void parse_tped(const std::string &tped, const std::string &tfam, Genotypes &out);
// Dispatch on extensions like Dataset<uint64_t>::read.
void parse_files(const std::vector<std::string> &paths, Genotypes &out);
// parse_files, then pack on the host exactly as the reference's readers do.
void read_files(const std::vector<std::string> &paths, PackedDataset &out);

}  // namespace fiu

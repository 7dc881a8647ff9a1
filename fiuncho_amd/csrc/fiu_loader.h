// Host-side genotype file readers (see fiu_loader.cpp).
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <stdexcept>
#include <string>
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

// Pack a SNP-major genotype matrix (values 0/1/2, anything else sets no bit).
void pack_genotypes(const std::vector<int8_t> &geno, const std::vector<uint8_t> &is_ctrl, size_t snps,
                    size_t samples, PackedDataset &out);
void read_raw(const std::string &path, PackedDataset &out);
void read_tped(const std::string &tped, const std::string &tfam, PackedDataset &out);
// Dispatch on extensions like Dataset<uint64_t>::read.
void read_files(const std::vector<std::string> &paths, PackedDataset &out);

}  // namespace fiu

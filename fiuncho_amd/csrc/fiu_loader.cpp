// Host-side readers of PLINK .raw / GAMETES .raw and .tped + .tfam files into
// the reference's bit-packed layout (ALIGN = 64 bytes).
//
// Behaviour follows the reference loaders (UDC-GAC/fiuncho):
//   RawFile::read_samples   include/fiuncho/dataset/RAWFile.hpp:149-207
//   RawFile::fill / read    RAWFile.hpp:209-313
//   TPEDFile::read_samples  include/fiuncho/dataset/TPEDFile.hpp:120-143
//   TPEDFile::read_variants TPEDFile.hpp:145-173
//   TPEDFile::fill / read   TPEDFile.hpp:175-296
//   Dataset::read dispatch  include/fiuncho/Dataset.h:48-94
// Samples whose phenotype is 1 are controls, every other value is a case.
// Within a class, samples keep file order; full 64-bit words hold samples
// most-significant-bit first and the last partial word is right-aligned;
// padding words and bits are zero.
#include "fiu_loader.h"

#include <cctype>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace fiu {
namespace {

size_t words_for(size_t n)
{
    // rows padded to whole 512-bit vectors (ALIGN = 64 bytes)
    return (n + 511) / 512 * 8;
}

std::string read_all(const std::string &path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw LoadError("Error while opening " + path + ", check file path/permissions");
    std::ostringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

std::vector<std::string> split_lines(const std::string &text)
{
    std::vector<std::string> lines;
    size_t pos = 0;
    while (pos < text.size()) {
        size_t e = text.find('\n', pos);
        if (e == std::string::npos) e = text.size();
        lines.emplace_back(text, pos, e - pos);
        pos = e + 1;
    }
    return lines;
}

bool is_word(char c) { return std::isalnum((unsigned char)c) || c == '_'; }

// A header word names a genotype column if it is X_[ACGT] or alphanumeric.
bool genotype_column(const std::string &w)
{
    bool alnum = !w.empty();
    for (char c : w) alnum = alnum && std::isalnum((unsigned char)c);
    if (alnum) return true;
    const size_t n = w.size();
    return n >= 3 && w[n - 2] == '_' && std::strchr("ACGT", w[n - 1]) != nullptr;
}

}  // namespace

void pack_genotypes(const std::vector<int8_t> &geno, const std::vector<uint8_t> &is_ctrl, size_t snps,
                    size_t samples, PackedDataset &out)
{
    size_t ctrls = 0;
    for (uint8_t c : is_ctrl) ctrls += c != 0;
    out.snps = snps;
    out.ctrls = ctrls;
    out.cases = samples - ctrls;
    out.cases_words = words_for(out.cases);
    out.ctrls_words = words_for(out.ctrls);
    const size_t per = 3 * (out.cases_words + out.ctrls_words);
    out.words.assign(snps * per, 0);
    for (size_t s = 0; s < snps; ++s) {
        const int8_t *g = geno.data() + s * samples;
        for (int cls = 0; cls < 2; ++cls) {
            const size_t W = cls ? out.ctrls_words : out.cases_words;
            uint64_t *rows = out.words.data() + s * per + (cls ? 3 * out.cases_words : 0);
            uint64_t acc[3] = {0, 0, 0};
            size_t n = 0;
            for (size_t j = 0; j < samples; ++j) {
                if ((is_ctrl[j] != 0) != (cls == 1)) continue;
                for (int z = 0; z < 3; ++z) acc[z] = (acc[z] << 1) | (uint64_t)(g[j] == z);
                if (++n % 64 == 0) {
                    for (int z = 0; z < 3; ++z) rows[z * W + n / 64 - 1] = acc[z], acc[z] = 0;
                }
            }
            if (n % 64)
                for (int z = 0; z < 3; ++z) rows[z * W + n / 64] = acc[z];
        }
    }
}

// The reference's fill loops (RAWFile.hpp:226-259, TPEDFile.hpp:213-245) test
// the "buffer full" condition of both classes after every sample, so once a
// class has 64k samples every following sample of the other class writes the
// cleared buffer into word k-1 again: that word (all three rows) is zero when
// the sample right after the 64k-th one belongs to the other class.  Files
// are read exactly as the reference reads them, so this is reproduced.
static void file_flush_quirk(const std::vector<uint8_t> &is_ctrl, PackedDataset &d)
{
    size_t cnt[2] = {0, 0};
    const size_t per = 3 * (d.cases_words + d.ctrls_words);
    for (size_t j = 0; j < is_ctrl.size(); ++j) {
        const int c = is_ctrl[j] != 0;
        ++cnt[c];
        if (cnt[c] % 64 == 0 && j + 1 < is_ctrl.size() && (is_ctrl[j + 1] != 0) != c) {
            const size_t W = c ? d.ctrls_words : d.cases_words;
            const size_t off = c ? 3 * d.cases_words : 0, w = cnt[c] / 64 - 1;
            for (size_t s = 0; s < d.snps; ++s)
                for (int k = 0; k < 3; ++k) d.words[s * per + off + k * W + w] = 0;
        }
    }
}

void read_raw(const std::string &path, PackedDataset &out)
{
    const std::string text = read_all(path);
    const std::vector<std::string> lines = split_lines(text);
    if (lines.empty()) throw LoadError("Error while parsing \"" + path + "\": empty file");
    // header words
    std::vector<std::string> words;
    const std::string &h = lines[0];
    for (size_t i = 0; i < h.size();) {
        if (!is_word(h[i])) { ++i; continue; }
        size_t j = i;
        while (j < h.size() && is_word(h[j])) ++j;
        words.emplace_back(h, i, j - i);
        i = j;
    }
    static const char *info[5] = {"FID", "IID", "PAT", "MAT", "SEX"};
    size_t ninfo = 0;
    while (ninfo < words.size()) {
        bool hit = false;
        for (const char *w : info) hit = hit || words[ninfo] == w;
        if (!hit) break;
        ++ninfo;
    }
    const bool has_info = ninfo == 5;
    const size_t first = has_info ? 5 : 0;
    std::vector<uint8_t> read_mask;
    long ph_idx = -1;
    size_t nvar = 0;
    for (size_t i = first; i < words.size(); ++i) {
        const std::string &w = words[i];
        if (w == "PHENOTYPE" || w == "Class") {
            if (ph_idx >= 0)
                throw LoadError("Error while parsing \"" + path + "\", multiple phenotype columns found");
            ph_idx = (long)(i - first);
            read_mask.push_back(1);
        } else {
            read_mask.push_back(genotype_column(w) ? 1 : 0);
            nvar += read_mask.back();
        }
    }
    std::vector<int8_t> by_sample;  // samples x nvar
    std::vector<uint8_t> is_ctrl;
    for (size_t li = 1; li < lines.size(); ++li) {
        const std::string &line = lines[li];
        if (li + 1 == lines.size() && line.empty()) break;
        const char *p = line.c_str();
        auto next_token = [&p]() -> std::string {
            while (*p == ' ' || *p == '\t' || *p == '\r') ++p;
            const char *b = p;
            while (*p && !std::isspace((unsigned char)*p)) ++p;
            return std::string(b, p - b);
        };
        auto next_short = [&](short &v) -> bool {
            const std::string t = next_token();
            if (t.empty()) return false;
            char *end = nullptr;
            const long x = std::strtol(t.c_str(), &end, 10);
            if (*end || x < -32768 || x > 32767) return false;
            v = (short)x;
            return true;
        };
        bool ok = true;
        if (has_info) {
            for (int q = 0; q < 4 && ok; ++q) ok = !next_token().empty();
            short sex;
            ok = ok && next_short(sex);
        }
        short ph = 0;
        for (size_t i = 0; i < read_mask.size() && ok; ++i) {
            short v;
            ok = next_short(v);
            if (!ok) break;
            if (read_mask[i]) {
                if ((long)i == ph_idx) ph = (short)(v + (has_info ? 0 : 1));
                else by_sample.push_back((int8_t)(v >= 0 && v <= 2 ? v : 3));
            }
        }
        if (!ok) throw LoadError("parsing error in \"" + path + "\" at line " + std::to_string(li + 1));
        is_ctrl.push_back(ph == 1 ? 1 : 0);
    }
    const size_t samples = is_ctrl.size();
    if (samples == 0) throw LoadError("Error while parsing \"" + path + "\": no samples");
    std::vector<int8_t> geno(nvar * samples);
    for (size_t j = 0; j < samples; ++j)
        for (size_t v = 0; v < nvar; ++v) geno[v * samples + j] = by_sample[j * nvar + v];
    pack_genotypes(geno, is_ctrl, nvar, samples, out);
    file_flush_quirk(is_ctrl, out);
}

void read_tped(const std::string &tped, const std::string &tfam, PackedDataset &out)
{
    std::vector<uint8_t> is_ctrl;
    {
        const std::vector<std::string> lines = split_lines(read_all(tfam));
        for (size_t li = 0; li < lines.size(); ++li) {
            if (li + 1 == lines.size() && lines[li].empty()) break;
            std::istringstream iss(lines[li]);
            std::string fid, iid, pat, mat;
            int sex, ph;
            if (!(iss >> fid >> iid >> pat >> mat >> sex >> ph) || (ph != 1 && ph != 2))
                throw LoadError("Error in " + tfam + ":" + std::to_string(li + 1) + ": parsing error");
            is_ctrl.push_back(ph == 1 ? 1 : 0);
        }
    }
    const size_t samples = is_ctrl.size();
    std::vector<int8_t> geno;
    size_t snps = 0;
    const std::vector<std::string> lines = split_lines(read_all(tped));
    std::vector<char> al;
    for (size_t li = 0; li < lines.size(); ++li) {
        const std::string &line = lines[li];
        if (li + 1 == lines.size() && line.empty()) break;
        std::istringstream iss(line);
        std::string chr, vid;
        double pos;
        unsigned coord;
        if (!(iss >> chr >> vid >> pos >> coord))
            throw LoadError("Error in " + tped + ":" + std::to_string(li + 1) + ": invalid loci information");
        al.clear();
        char c;
        while (iss >> c) {
            if (!std::strchr("ACGT", c))
                throw LoadError("Error in " + tped + ":" + std::to_string(li + 1) +
                                ": invalid nucleotide value '" + std::string(1, c) + "'");
            al.push_back(c);
        }
        if (al.size() % 2) throw LoadError("Error in " + tped + ":" + std::to_string(li + 1) + ": odd number of nucleotides");
        if (al.size() != 2 * samples)
            throw LoadError("Error in " + tped + ":" + std::to_string(li + 1) +
                            ": the number of nucleotides does not match the number of samples");
        // minor allele: the less frequent of the two alphabetically first
        // alleles present, ties to the first; none on a monomorphic SNP
        size_t cnt[4] = {0, 0, 0, 0};
        static const char order[4] = {'A', 'C', 'G', 'T'};
        for (char a : al)
            for (int q = 0; q < 4; ++q) cnt[q] += a == order[q];
        int a1 = -1, a2 = -1;
        for (int q = 0; q < 4; ++q)
            if (cnt[q]) {
                if (a1 < 0) a1 = q;
                else if (a2 < 0) a2 = q;
            }
        char minor = 0;
        if (a2 >= 0) minor = cnt[a1] > cnt[a2] ? order[a2] : order[a1];
        for (size_t j = 0; j < samples; ++j) geno.push_back((int8_t)((al[2 * j] == minor) + (al[2 * j + 1] == minor)));
        ++snps;
    }
    pack_genotypes(geno, is_ctrl, snps, samples, out);
    file_flush_quirk(is_ctrl, out);
}

void read_files(const std::vector<std::string> &paths, PackedDataset &out)
{
    auto ext = [](const std::string &p) {
        const size_t slash = p.find_last_of('/');
        const size_t dot = p.find_last_of('.');
        if (dot == std::string::npos || (slash != std::string::npos && dot < slash)) return std::string();
        return p.substr(dot);
    };
    if (paths.size() == 1) {
        if (ext(paths[0]) != ".raw") throw LoadError("Unrecognized input file extension " + ext(paths[0]));
        read_raw(paths[0], out);
    } else if (paths.size() == 2) {
        const std::string e1 = ext(paths[0]), e2 = ext(paths[1]);
        if (e1 != ".tped" && e1 != ".tfam") throw LoadError("Unrecognized input file extension " + e1);
        if (e1 == e2 || (e2 != ".tped" && e2 != ".tfam"))
            throw LoadError("Expecting a tfam file after \"" + paths[0] + "\", but found a " + e2);
        if (e1 == ".tped") read_tped(paths[0], paths[1], out);
        else read_tped(paths[1], paths[0], out);
    } else {
        throw LoadError("Unrecognized number of input files to Dataset factory constructor");
    }
}

}  // namespace fiu

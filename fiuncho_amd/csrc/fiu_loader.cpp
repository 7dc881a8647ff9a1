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

#include <algorithm>
#include <atomic>
#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <thread>

namespace fiu {
namespace {

std::string read_all(const std::string &path)
{
    std::FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) throw LoadError("Error while opening " + path + ", check file path/permissions");
    std::string text;
    char buf[1 << 16];
    size_t got;
    if (std::fseek(f, 0, SEEK_END) == 0) {
        const long size = std::ftell(f);
        if (size > 0) text.reserve((size_t)size);
        std::rewind(f);
    }
    while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, got);
    const bool err = std::ferror(f) != 0;
    std::fclose(f);
    if (err) throw LoadError("Error while reading " + path);
    return text;
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
    out.cases_words = packed_words(out.cases);
    out.ctrls_words = packed_words(out.ctrls);
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
std::vector<std::pair<int, size_t>> flush_quirk_words(const uint8_t *is_ctrl, size_t samples)
{
    std::vector<std::pair<int, size_t>> out;
    size_t cnt[2] = {0, 0};
    for (size_t j = 0; j < samples; ++j) {
        const int c = is_ctrl[j] != 0;
        ++cnt[c];
        if (cnt[c] % 64 == 0 && j + 1 < samples && (is_ctrl[j + 1] != 0) != c)
            out.emplace_back(c, cnt[c] / 64 - 1);
    }
    return out;
}

static void file_flush_quirk(const std::vector<uint8_t> &is_ctrl, PackedDataset &d)
{
    const size_t per = 3 * (d.cases_words + d.ctrls_words);
    for (const auto &cw : flush_quirk_words(is_ctrl.data(), is_ctrl.size())) {
        const int c = cw.first;
        const size_t W = c ? d.ctrls_words : d.cases_words, off = c ? 3 * d.cases_words : 0, w = cw.second;
        for (size_t s = 0; s < d.snps; ++s)
            for (int k = 0; k < 3; ++k) d.words[s * per + off + k * W + w] = 0;
    }
}

namespace {

struct Span {
    const char *p;
    size_t n;
};

// Lines of `text` split at '\n' (a last line without '\n' counts; a final
// '\n' does not open an empty line).
std::vector<Span> line_spans(const std::string &text)
{
    std::vector<Span> lines;
    const char *b = text.data(), *e = b + text.size();
    while (b < e) {
        const char *nl = (const char *)std::memchr(b, '\n', (size_t)(e - b));
        const char *le = nl ? nl : e;
        lines.push_back({b, (size_t)(le - b)});
        b = le + 1;
    }
    return lines;
}

inline bool space(char c) { return c == ' ' || (c >= '\t' && c <= '\r'); }

// Next token of a sample line: skip ' ', '\t', '\r', then a run of
// non-space characters.
inline Span next_token(const char *&p, const char *e)
{
    while (p < e && (*p == ' ' || *p == '\t' || *p == '\r')) ++p;
    const char *b = p;
    while (p < e && !space(*p)) ++p;
    return {b, (size_t)(p - b)};
}

// A token that is a whole base-10 integer (optional sign) within short.
inline bool next_short(const char *&p, const char *e, short &v)
{
    const Span t = next_token(p, e);
    const char *q = t.p, *end = t.p + t.n;
    if (q == end) return false;
    const bool neg = *q == '-';
    if (*q == '-' || *q == '+') ++q;
    if (q == end) return false;
    long x = 0;
    for (; q < end; ++q) {
        const unsigned d = (unsigned)(unsigned char)*q - '0';
        if (d > 9) return false;
        x = x * 10 + d;
        if (x > 32768) return false;
    }
    x = neg ? -x : x;
    if (x < -32768 || x > 32767) return false;
    v = (short)x;
    return true;
}

int loader_threads(size_t items)
{
    const unsigned hw = std::max(1u, std::min(std::thread::hardware_concurrency(), 16u));
    return (int)std::max<size_t>(1, std::min<size_t>(hw, items / 64));
}

// f(lo, hi) over contiguous parts of [0, n) on up to `threads` threads.
template <class F> void parallel_ranges(size_t n, int threads, const F &f)
{
    if (threads <= 1 || n < 2) {
        f((size_t)0, n);
        return;
    }
    std::vector<std::thread> pool;
    const size_t step = (n + threads - 1) / threads;
    for (size_t lo = 0; lo < n; lo += step) pool.emplace_back(f, lo, std::min(n, lo + step));
    for (auto &t : pool) t.join();
}

}  // namespace

void parse_raw(const std::string &path, Genotypes &out)
{
    const std::string text = read_all(path);
    const std::vector<Span> lines = line_spans(text);
    if (lines.empty()) throw LoadError("Error while parsing \"" + path + "\": empty file");
    // header words
    std::vector<std::string> words;
    const std::string h(lines[0].p, lines[0].n);
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
    // one sample per line after the header; a last empty line ends the file
    size_t samples = lines.size() - 1;
    if (samples && lines.back().n == 0) --samples;
    if (samples == 0) throw LoadError("Error while parsing \"" + path + "\": no samples");
    std::vector<int8_t> by_sample(samples * nvar);   // samples x nvar
    std::vector<uint8_t> is_ctrl(samples);
    std::vector<size_t> bad(std::max(loader_threads(samples), 1), SIZE_MAX);
    std::atomic<int> part{0};
    parallel_ranges(samples, loader_threads(samples), [&](size_t lo, size_t hi) {
        const int me = part.fetch_add(1);
        for (size_t j = lo; j < hi; ++j) {
            const char *p = lines[j + 1].p, *e = p + lines[j + 1].n;
            int8_t *dst = by_sample.data() + j * nvar;
            bool ok = true;
            if (has_info) {
                for (int q = 0; q < 4 && ok; ++q) ok = next_token(p, e).n != 0;
                short sex;
                ok = ok && next_short(p, e, sex);
            }
            short ph = 0;
            for (size_t i = 0; i < read_mask.size() && ok; ++i) {
                short v;
                ok = next_short(p, e, v);
                if (!ok) break;
                if (read_mask[i]) {
                    if ((long)i == ph_idx) ph = (short)(v + (has_info ? 0 : 1));
                    else *dst++ = (int8_t)(v >= 0 && v <= 2 ? v : 3);
                }
            }
            if (!ok) {
                bad[me] = j;   // the first failing line of this part
                return;
            }
            is_ctrl[j] = ph == 1 ? 1 : 0;
        }
    });
    const size_t first_bad = *std::min_element(bad.begin(), bad.end());
    if (first_bad != SIZE_MAX)
        throw LoadError("parsing error in \"" + path + "\" at line " + std::to_string(first_bad + 2));
    // SNP-major, in 64 x 64 tiles
    std::vector<int8_t> &geno = out.geno;
    geno.assign(nvar * samples, 0);
    const size_t tiles = (samples + 63) / 64;
    parallel_ranges(tiles, loader_threads(tiles * 64), [&](size_t lo, size_t hi) {
        for (size_t t = lo; t < hi; ++t)
            for (size_t v0 = 0; v0 < nvar; v0 += 64)
                for (size_t j = t * 64; j < std::min(samples, t * 64 + 64); ++j)
                    for (size_t v = v0; v < std::min(nvar, v0 + 64); ++v)
                        geno[v * samples + j] = by_sample[j * nvar + v];
    });
    out.snps = nvar;
    out.samples = samples;
    out.is_ctrl = std::move(is_ctrl);
}

void parse_tped(const std::string &tped, const std::string &tfam, Genotypes &out)
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
    const std::string text = read_all(tped);
    const std::vector<Span> lines = line_spans(text);
    size_t snps = lines.size();
    if (snps && lines.back().n == 0) --snps;   // a last empty line ends the file
    std::vector<int8_t> geno(snps * samples);
    // one SNP per line, lines in parallel; each part keeps its first error
    const int threads = loader_threads(snps);
    std::vector<std::pair<size_t, std::string>> bad(std::max(threads, 1), {SIZE_MAX, std::string()});
    std::atomic<int> part{0};
    parallel_ranges(snps, threads, [&](size_t lo, size_t hi) {
        const int me = part.fetch_add(1);
        std::vector<uint8_t> al;   // allele codes 0..3 = A, C, G, T
        al.reserve(2 * samples);
        for (size_t li = lo; li < hi; ++li) {
            auto err = [&](const std::string &what) { bad[me] = {li, "Error in " + tped + ":" + std::to_string(li + 1) + ": " + what}; };
            const std::string line(lines[li].p, lines[li].n);
            std::istringstream iss(line);
            std::string chr, vid;
            double pos;
            unsigned coord;
            if (!(iss >> chr >> vid >> pos >> coord)) return err("invalid loci information");
            const std::streamoff at = iss.tellg();
            const char *p = line.data() + (at < 0 ? line.size() : (size_t)at), *e = line.data() + line.size();
            al.clear();
            for (; p < e; ++p) {
                const char c = *p;
                if (space(c)) continue;
                static const char kAlleles[] = "ACGT";
                const char *q = c ? std::strchr(kAlleles, c) : nullptr;
                if (!q) return err("invalid nucleotide value '" + std::string(1, c) + "'");
                al.push_back((uint8_t)(q - kAlleles));
            }
            if (al.size() % 2) return err("odd number of nucleotides");
            if (al.size() != 2 * samples) return err("the number of nucleotides does not match the number of samples");
            // minor allele: the less frequent of the two alphabetically
            // first alleles present, ties to the first; none on a
            // monomorphic SNP
            size_t cnt[4] = {0, 0, 0, 0};
            for (uint8_t a : al) ++cnt[a];
            int a1 = -1, a2 = -1;
            for (int q = 0; q < 4; ++q)
                if (cnt[q]) {
                    if (a1 < 0) a1 = q;
                    else if (a2 < 0) a2 = q;
                }
            const int minor = a2 < 0 ? -1 : cnt[a1] > cnt[a2] ? a2 : a1;
            int8_t *g = geno.data() + li * samples;
            for (size_t j = 0; j < samples; ++j) g[j] = (int8_t)((al[2 * j] == minor) + (al[2 * j + 1] == minor));
        }
    });
    const auto first_bad = std::min_element(bad.begin(), bad.end(),
                                            [](const auto &x, const auto &y) { return x.first < y.first; });
    if (first_bad->first != SIZE_MAX) throw LoadError(first_bad->second);
    out.geno = std::move(geno);
    out.snps = snps;
    out.samples = samples;
    out.is_ctrl = std::move(is_ctrl);
}

void parse_files(const std::vector<std::string> &paths, Genotypes &out)
{
    auto ext = [](const std::string &p) {
        const size_t slash = p.find_last_of('/');
        const size_t dot = p.find_last_of('.');
        if (dot == std::string::npos || (slash != std::string::npos && dot < slash)) return std::string();
        return p.substr(dot);
    };
    if (paths.size() == 1) {
        if (ext(paths[0]) != ".raw") throw LoadError("Unrecognized input file extension " + ext(paths[0]));
        parse_raw(paths[0], out);
    } else if (paths.size() == 2) {
        const std::string e1 = ext(paths[0]), e2 = ext(paths[1]);
        if (e1 != ".tped" && e1 != ".tfam") throw LoadError("Unrecognized input file extension " + e1);
        if (e1 == e2 || (e2 != ".tped" && e2 != ".tfam"))
            throw LoadError("Expecting a tfam file after \"" + paths[0] + "\", but found a " + e2);
        if (e1 == ".tped") parse_tped(paths[0], paths[1], out);
        else parse_tped(paths[1], paths[0], out);
    } else {
        throw LoadError("Unrecognized number of input files to Dataset factory constructor");
    }
}

void read_files(const std::vector<std::string> &paths, PackedDataset &out)
{
    Genotypes g;
    parse_files(paths, g);
    pack_genotypes(g.geno, g.is_ctrl, g.snps, g.samples, out);
    file_flush_quirk(g.is_ctrl, out);
}

}  // namespace fiu

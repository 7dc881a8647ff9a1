/*
 * CPU oracle for the exhaustive epistasis search -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, the smoke check and the benchmark's CPU-baseline leg load this
 * library, as the checker.  The product (fiuncho_amd/) never links it.
 *
 * Plain-C restatement of the reference algorithm (UDC-GAC/fiuncho); each
 * function cites the reference file:line it follows.  Pinned by
 * tests/test_oracle.py (reference fixtures, known answers, brute force and
 * vectors from the reference compiled from source).
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <ctype.h>
#include <errno.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#if defined(__x86_64__)
#include <immintrin.h>
#endif

/* ------------------------------------------------------------------------ */
/* Layout helpers                                                            */
/* ------------------------------------------------------------------------ */

size_t orc_words_for(size_t n, int align)
{
    /* RAWFile.hpp:299-309: with ALIGN bytes, rows are padded to a multiple of
     * ALIGN*8 bits, i.e. ceil(n / (ALIGN*8)) * (ALIGN/8) words. */
    if (align <= 0) return (n + 63) / 64;
    size_t nbits = (size_t)align * 8, nt = (size_t)align / 8;
    return (n + nbits - 1) / nbits * nt;
}

size_t orc_ct_size(int order, int align)
{
    /* ContingencyTable.h:53-65: 3^order rounded up to ALIGN/4 entries. */
    size_t s = 1;
    for (int i = 0; i < order; ++i) s *= 3;
    if (align <= 0) return s;
    size_t nt = (size_t)align / 4;
    return (s + nt - 1) / nt * nt;
}

static void set_err(char *err, size_t errlen, const char *fmt, const char *a,
                    long b)
{
    if (err && errlen) snprintf(err, errlen, fmt, a, b);
}

/* Pack bits for one SNP and one class, following the fill loops of
 * RAWFile.hpp:209-290 and TPEDFile.hpp:175-269: samples of the class in file
 * order, full words MSB-first (buff = buff<<1 | bit, flushed every 64), the
 * last partial word right-aligned, padding words zero. */
static void pack_class_rows(const int8_t *g, const uint8_t *is_ctrl,
                            size_t samples, int want_ctrl, uint64_t *rows,
                            size_t words)
{
    memset(rows, 0, 3 * words * sizeof(uint64_t));
    uint64_t buf[3] = {0, 0, 0};
    size_t cnt = 0;
    for (size_t j = 0; j < samples; ++j) {
        if ((is_ctrl[j] != 0) != (want_ctrl != 0)) continue;
        for (int k = 0; k < 3; ++k) buf[k] = (buf[k] << 1) | (uint64_t)(g[j] == k);
        ++cnt;
        if (cnt % 64 == 0) {
            for (int k = 0; k < 3; ++k) {
                rows[k * words + cnt / 64 - 1] = buf[k];
                buf[k] = 0;
            }
        }
    }
    if (cnt % 64)
        for (int k = 0; k < 3; ++k) rows[k * words + cnt / 64] = buf[k];
}

/* The fill loops of RAWFile.hpp:226-259 / TPEDFile.hpp:213-245 test the
 * "buffer full" condition of BOTH classes after every sample, so after the
 * 64k-th sample of a class every following sample of the other class flushes
 * the (already cleared) buffer again into word k-1: that word ends up zero
 * when the sample right after the 64k-th one belongs to the other class.
 * Applied to datasets read from files, as the reference reads them. */
static void file_flush_quirk(const uint8_t *is_ctrl, size_t samples,
                             orc_dataset *d)
{
    size_t cnt[2] = {0, 0};
    const size_t per = 3 * (d->cases_words + d->ctrls_words);
    for (size_t j = 0; j < samples; ++j) {
        const int c = is_ctrl[j] != 0;
        ++cnt[c];
        if (cnt[c] % 64 == 0 && j + 1 < samples && (is_ctrl[j + 1] != 0) != c) {
            const size_t W = c ? d->ctrls_words : d->cases_words;
            const size_t off = c ? 3 * d->cases_words : 0, w = cnt[c] / 64 - 1;
            for (size_t s = 0; s < d->snps; ++s)
                for (int k = 0; k < 3; ++k) d->words[s * per + off + k * W + w] = 0;
        }
    }
}

int orc_pack(const int8_t *geno, const uint8_t *is_ctrl, size_t snps,
             size_t samples, int align, orc_dataset *out)
{
    size_t ctrls = 0;
    for (size_t j = 0; j < samples; ++j) ctrls += is_ctrl[j] != 0;
    out->snps = snps;
    out->ctrls = ctrls;
    out->cases = samples - ctrls;
    out->cases_words = orc_words_for(out->cases, align);
    out->ctrls_words = orc_words_for(out->ctrls, align);
    size_t per = 3 * (out->cases_words + out->ctrls_words);
    out->words = (uint64_t *)calloc(snps * per + 1, sizeof(uint64_t));
    if (!out->words) return -1;
    for (size_t s = 0; s < snps; ++s) {
        uint64_t *base = out->words + s * per;
        pack_class_rows(geno + s * samples, is_ctrl, samples, 0, base,
                        out->cases_words);
        pack_class_rows(geno + s * samples, is_ctrl, samples, 1,
                        base + 3 * out->cases_words, out->ctrls_words);
    }
    return 0;
}

void orc_dataset_free(orc_dataset *d)
{
    if (d && d->words) free(d->words);
    if (d) d->words = NULL;
}

/* ------------------------------------------------------------------------ */
/* Loaders                                                                   */
/* ------------------------------------------------------------------------ */

static char *read_file(const char *path, size_t *len)
{
    FILE *f = fopen(path, "rb");
    if (!f) return NULL;
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    char *buf = (char *)malloc((size_t)n + 1);
    if (!buf) { fclose(f); return NULL; }
    size_t got = fread(buf, 1, (size_t)n, f);
    fclose(f);
    buf[got] = 0;
    *len = got;
    return buf;
}

static int is_word_char(int c) { return isalnum(c) || c == '_'; }

/* Header column classification of RAWFile.hpp:166-196: a column is read as a
 * genotype if its name matches (\w+_[ACGT]) or ([[:alnum:]]+). */
static int raw_genotype_column(const char *w, size_t n)
{
    int alnum = 1;
    for (size_t i = 0; i < n; ++i)
        if (!isalnum((unsigned char)w[i])) alnum = 0;
    if (alnum && n > 0) return 1;
    if (n >= 3 && w[n - 2] == '_' && strchr("ACGT", w[n - 1])) return 1;
    return 0;
}

/* Parse one whitespace-delimited integer like `istream >> short`. */
static int parse_short(const char **p, short *v)
{
    const char *s = *p;
    while (*s == ' ' || *s == '\t' || *s == '\r') ++s;
    if (!*s || *s == '\n') return 0;
    char *end;
    errno = 0;
    long x = strtol(s, &end, 10);
    if (end == s || errno || x < -32768 || x > 32767) return 0;
    *p = end;
    *v = (short)x;
    return 1;
}

int orc_read_raw(const char *path, int align, orc_dataset *out, char *err,
                 size_t errlen)
{
    /* RAWFile.hpp:149-207 (read_samples) and :292-313 (read). */
    size_t len;
    char *buf = read_file(path, &len);
    if (!buf) { set_err(err, errlen, "cannot open %s%ld", path, 0); return -1; }
    char *line = buf, *nl = strchr(line, '\n');
    if (nl) *nl = 0;
    /* tokenize header into \w+ words */
    size_t cap = 64, ntok = 0;
    char **tok = (char **)malloc(cap * sizeof(char *));
    size_t *tlen = (size_t *)malloc(cap * sizeof(size_t));
    for (char *s = line; *s;) {
        if (!is_word_char((unsigned char)*s)) { ++s; continue; }
        char *b = s;
        while (*s && is_word_char((unsigned char)*s)) ++s;
        if (ntok == cap) {
            cap *= 2;
            tok = (char **)realloc(tok, cap * sizeof(char *));
            tlen = (size_t *)realloc(tlen, cap * sizeof(size_t));
        }
        tok[ntok] = b;
        tlen[ntok++] = (size_t)(s - b);
    }
    static const char *info[5] = {"FID", "IID", "PAT", "MAT", "SEX"};
    size_t ninfo = 0;
    while (ninfo < ntok) {
        int hit = 0;
        for (int q = 0; q < 5; ++q)
            if (tlen[ninfo] == 3 && !strncmp(tok[ninfo], info[q], 3)) hit = 1;
        if (!hit) break;
        ++ninfo;
    }
    int has_info = ninfo == 5;
    size_t first = has_info ? 5 : 0, ncol = ntok - first;
    uint8_t *mask = (uint8_t *)calloc(ncol + 1, 1);
    long ph_idx = -1;
    size_t nvar = 0;
    for (size_t i = 0; i < ncol; ++i) {
        const char *w = tok[first + i];
        size_t n = tlen[first + i];
        if ((n == 9 && !strncmp(w, "PHENOTYPE", 9)) ||
            (n == 5 && !strncmp(w, "Class", 5))) {
            if (ph_idx >= 0) {
                set_err(err, errlen, "multiple phenotype columns in %s%ld",
                        path, 0);
                free(buf); free(tok); free(tlen); free(mask);
                return -1;
            }
            ph_idx = (long)i;
            mask[i] = 1;
        } else {
            mask[i] = (uint8_t)raw_genotype_column(w, n);
            nvar += mask[i];
        }
    }
    free(tok);
    free(tlen);
    /* samples */
    size_t scap = 1024, ns = 0;
    int8_t *g = NULL; /* sample-major while reading: ns x nvar */
    short *ph = (short *)malloc(scap * sizeof(short));
    g = (int8_t *)malloc(scap * (nvar ? nvar : 1));
    char *s = nl ? nl + 1 : buf + len;
    while (s < buf + len) {
        char *e = strchr(s, '\n');
        if (e) *e = 0;
        char *next = e ? e + 1 : buf + len;
        /* std::getline succeeds on any line, even an empty trailing one
         * before EOF is not produced; skip a final empty line at EOF. */
        if (!e && !*s) break;
        const char *p = s;
        if (ns == scap) {
            scap *= 2;
            ph = (short *)realloc(ph, scap * sizeof(short));
            g = (int8_t *)realloc(g, scap * (nvar ? nvar : 1));
        }
        int ok = 1;
        if (has_info) {
            for (int q = 0; q < 4 && ok; ++q) {
                while (*p == ' ' || *p == '\t' || *p == '\r') ++p;
                if (!*p) ok = 0;
                while (*p && !isspace((unsigned char)*p)) ++p;
            }
            short sex;
            if (ok) ok = parse_short(&p, &sex);
        }
        size_t v = 0;
        short phv = 0;
        for (size_t i = 0; i < ncol && ok; ++i) {
            short val;
            ok = parse_short(&p, &val);
            if (!ok) break;
            if (mask[i]) {
                if ((long)i == ph_idx) phv = (short)(val + !has_info);
                else g[ns * nvar + v++] = (int8_t)(val >= 0 && val <= 2 ? val : 3);
            }
        }
        if (!ok) {
            set_err(err, errlen, "parsing error in %s at sample %ld", path,
                    (long)ns);
            free(buf); free(mask); free(ph); free(g);
            return -1;
        }
        ph[ns++] = phv;
        s = next;
    }
    free(buf);
    free(mask);
    /* RAWFile.hpp:203,230: phenotype 1 is a control, anything else a case */
    int8_t *gm = (int8_t *)malloc(nvar * ns + 1);
    uint8_t *ctrl = (uint8_t *)malloc(ns + 1);
    for (size_t j = 0; j < ns; ++j) {
        ctrl[j] = ph[j] == 1;
        for (size_t i = 0; i < nvar; ++i) gm[i * ns + j] = g[j * nvar + i];
    }
    free(ph);
    free(g);
    int rc = orc_pack(gm, ctrl, nvar, ns, align, out);
    if (rc == 0) file_flush_quirk(ctrl, ns, out);
    free(gm);
    free(ctrl);
    return rc;
}

int orc_read_tped(const char *tped, const char *tfam, int align,
                  orc_dataset *out, char *err, size_t errlen)
{
    /* TFAM: TPEDFile.hpp:87-143, phenotype must be 1 or 2. */
    size_t len;
    char *buf = read_file(tfam, &len);
    if (!buf) { set_err(err, errlen, "cannot open %s%ld", tfam, 0); return -1; }
    size_t cap = 1024, ns = 0;
    uint8_t *ctrl = (uint8_t *)malloc(cap);
    for (char *s = buf; s < buf + len;) {
        char *e = strchr(s, '\n');
        if (e) *e = 0;
        char *next = e ? e + 1 : buf + len;
        if (!e && !*s) break;
        char f[4][256];
        int sex, phv;
        if (sscanf(s, "%255s %255s %255s %255s %d %d", f[0], f[1], f[2], f[3],
                   &sex, &phv) != 6 || (phv != 1 && phv != 2)) {
            set_err(err, errlen, "parsing error in %s line %ld", tfam,
                    (long)ns + 1);
            free(buf); free(ctrl);
            return -1;
        }
        if (ns == cap) { cap *= 2; ctrl = (uint8_t *)realloc(ctrl, cap); }
        ctrl[ns++] = phv == 1;
        s = next;
    }
    free(buf);
    /* TPED: TPEDFile.hpp:28-61 (variant parse), :145-173 (count check),
     * :185-208 (minor allele = lower count, ties -> alphabetically first). */
    buf = read_file(tped, &len);
    if (!buf) { set_err(err, errlen, "cannot open %s%ld", tped, 0); free(ctrl); return -1; }
    size_t vcap = 64, nv = 0;
    int8_t *gm = (int8_t *)malloc(vcap * ns + 1);
    char *al = (char *)malloc(2 * ns + 2);
    for (char *s = buf; s < buf + len;) {
        char *e = strchr(s, '\n');
        if (e) *e = 0;
        char *next = e ? e + 1 : buf + len;
        if (!e && !*s) break;
        char chr[256], vid[256];
        double pos;
        unsigned coord;
        int consumed = 0;
        if (sscanf(s, "%255s %255s %lf %u%n", chr, vid, &pos, &coord,
                   &consumed) != 4) {
            set_err(err, errlen, "invalid loci information in %s line %ld",
                    tped, (long)nv + 1);
            free(buf); free(ctrl); free(gm); free(al);
            return -1;
        }
        size_t na = 0;
        for (const char *p = s + consumed; *p; ++p) {
            if (isspace((unsigned char)*p)) continue;
            if (!strchr("ACGT", *p)) {
                set_err(err, errlen, "invalid nucleotide in %s line %ld",
                        tped, (long)nv + 1);
                free(buf); free(ctrl); free(gm); free(al);
                return -1;
            }
            if (na < 2 * ns + 1) al[na] = *p;
            ++na;
        }
        if (na % 2 || na != 2 * ns) {
            set_err(err, errlen, "nucleotide count mismatch in %s line %ld",
                    tped, (long)nv + 1);
            free(buf); free(ctrl); free(gm); free(al);
            return -1;
        }
        size_t cnt[4] = {0, 0, 0, 0};
        static const char order[4] = {'A', 'C', 'G', 'T'};
        for (size_t j = 0; j < na; ++j)
            for (int q = 0; q < 4; ++q) cnt[q] += al[j] == order[q];
        int a1 = -1, a2 = -1;
        for (int q = 0; q < 4; ++q)
            if (cnt[q]) { if (a1 < 0) a1 = q; else if (a2 < 0) a2 = q; }
        /* Monomorphic SNP: the minor allele is absent, every genotype is 0. */
        char minor = 0;
        if (a2 >= 0) minor = cnt[a1] > cnt[a2] ? order[a2] : order[a1];
        if (nv == vcap) { vcap *= 2; gm = (int8_t *)realloc(gm, vcap * ns + 1); }
        for (size_t j = 0; j < ns; ++j)
            gm[nv * ns + j] = (int8_t)((al[2 * j] == minor) + (al[2 * j + 1] == minor));
        ++nv;
        s = next;
    }
    free(buf);
    free(al);
    int rc = orc_pack(gm, ctrl, nv, ns, align, out);
    if (rc == 0) file_flush_quirk(ctrl, ns, out);
    free(gm);
    free(ctrl);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Table operators                                                           */
/* ------------------------------------------------------------------------ */

void orc_combine(const uint64_t *t1, size_t size1, const uint64_t *t2,
                 size_t words, uint64_t *out)
{
    /* src/base/gt_combine.cpp:3-18: out row i*3+j = t1 row i AND t2 row j */
    for (size_t i = 0; i < size1; ++i)
        for (size_t j = 0; j < 3; ++j)
            for (size_t w = 0; w < words; ++w)
                out[(i * 3 + j) * words + w] = t1[i * words + w] & t2[j * words + w];
}

void orc_popcnt(const uint64_t *t1, size_t size1, const uint64_t *t2,
                size_t words, uint32_t *ct, size_t ct_size)
{
    /* src/base/gt_popcnt.cpp:4-23 */
    for (size_t i = 0; i < size1; ++i)
        for (size_t j = 0; j < 3; ++j) {
            uint32_t c = 0;
            for (size_t w = 0; w < words; ++w)
                c += (uint32_t)__builtin_popcountll(t1[i * words + w] & t2[j * words + w]);
            ct[i * 3 + j] = c;
        }
    for (size_t i = size1 * 3; i < ct_size; ++i) ct[i] = 0;
}

static const uint64_t *snp_cases(const orc_dataset *d, int s)
{
    return d->words + (size_t)s * 3 * (d->cases_words + d->ctrls_words);
}
static const uint64_t *snp_ctrls(const orc_dataset *d, int s)
{
    return snp_cases(d, s) + 3 * d->cases_words;
}

void orc_counts(const orc_dataset *d, const int *combo, int k,
                uint32_t *cases_ct, uint32_t *ctrls_ct)
{
    /* prefix chain of ThreadedSearch.h:160-164 then the last popcount of
     * :174-176 */
    size_t cw = d->cases_words, tw = d->ctrls_words, size = 3;
    for (int i = 1; i < k - 1; ++i) size *= 3;
    uint64_t *pc = (uint64_t *)malloc(size * cw * sizeof(uint64_t) + 8);
    uint64_t *pt = (uint64_t *)malloc(size * tw * sizeof(uint64_t) + 8);
    uint64_t *qc = (uint64_t *)malloc(size * cw * sizeof(uint64_t) + 8);
    uint64_t *qt = (uint64_t *)malloc(size * tw * sizeof(uint64_t) + 8);
    memcpy(pc, snp_cases(d, combo[0]), 3 * cw * sizeof(uint64_t));
    memcpy(pt, snp_ctrls(d, combo[0]), 3 * tw * sizeof(uint64_t));
    size_t cur = 3;
    for (int i = 1; i < k - 1; ++i) {
        orc_combine(pc, cur, snp_cases(d, combo[i]), cw, qc);
        orc_combine(pt, cur, snp_ctrls(d, combo[i]), tw, qt);
        cur *= 3;
        uint64_t *t = pc; pc = qc; qc = t;
        t = pt; pt = qt; qt = t;
    }
    orc_popcnt(pc, cur, snp_cases(d, combo[k - 1]), cw, cases_ct, cur * 3);
    orc_popcnt(pt, cur, snp_ctrls(d, combo[k - 1]), tw, ctrls_ct, cur * 3);
    free(pc); free(pt); free(qc); free(qt);
}

void orc_counts_bruteforce(const int8_t *geno, const uint8_t *is_ctrl,
                           size_t samples, const int *combo, int k,
                           uint32_t *cases_ct, uint32_t *ctrls_ct)
{
    /* Closed form: cell = sum_q g_q * 3^(k-1-q) (first SNP most significant,
     * from the i*3+j indexing of src/base/gt_combine.cpp:12-14); samples with
     * a genotype outside {0,1,2} in any SNP fall in no cell. */
    size_t cells = 1;
    for (int q = 0; q < k; ++q) cells *= 3;
    memset(cases_ct, 0, cells * sizeof(uint32_t));
    memset(ctrls_ct, 0, cells * sizeof(uint32_t));
    for (size_t j = 0; j < samples; ++j) {
        size_t cell = 0;
        int ok = 1;
        for (int q = 0; q < k; ++q) {
            int g = geno[(size_t)combo[q] * samples + j];
            if (g < 0 || g > 2) ok = 0;
            cell = cell * 3 + (size_t)g;
        }
        if (!ok) continue;
        if (is_ctrl[j]) ++ctrls_ct[cell];
        else ++cases_ct[cell];
    }
}

/* ------------------------------------------------------------------------ */
/* Mutual information                                                        */
/* ------------------------------------------------------------------------ */

/* MutualInformation.h:58-65: inv_inds is 1/N rounded to float; H(Y) takes
 * its first term in double and the second as a float multiply-subtract
 * (fused, as GCC contracts it in the reference's GNU C++ build). */
static void mi_consts(unsigned cases, unsigned ctrls, float *inv, float *hy)
{
    float inv_inds = (float)(1.0 / (double)(cases + ctrls));
    float p = (float)cases * inv_inds;
    float h = (float)((-1.0 * (double)p) * (double)logf(p));
    p = (float)ctrls * inv_inds;
    h = fmaf(-p, logf(p), h);
    *inv = inv_inds;
    *hy = h;
}

float orc_mi_f32_base(const uint32_t *cases_ct, const uint32_t *ctrls_ct,
                      size_t size, unsigned cases, unsigned ctrls)
{
    /* src/base/mi_compute.cpp:14-41, sequential over cells, zero cells
     * skipped; the multiply-subtracts fused as in the reference build. */
    float inv, hy;
    mi_consts(cases, ctrls, &inv, &hy);
    float hx = 0.0f, hall = 0.0f;
    for (size_t i = 0; i < size; ++i) {
        float pc = (float)cases_ct[i] * inv;
        if (pc != 0.0f) hall = fmaf(-pc, logf(pc), hall);
        float pt = (float)ctrls_ct[i] * inv;
        if (pt != 0.0f) hall = fmaf(-pt, logf(pt), hall);
        float pa = pc + pt;
        if (pa != 0.0f) hx = fmaf(-pa, logf(pa), hx);
    }
    return hx + hy - hall;
}

#if defined(__x86_64__)
/* glibc libmvec 16-lane logf (the function src/avx512bw/mi_compute_512.cpp
 * binds at :13-18). */
extern __m512 _ZGVeN16v_logf(__m512 x);

__attribute__((target("avx512f,avx512bw,avx512dq,fma"))) static float
mi_ref512_impl(const uint32_t *cases_ct, const uint32_t *ctrls_ct, size_t size,
               unsigned cases, unsigned ctrls)
{
    /* src/avx512bw/mi_compute_512.cpp:20-69: 16-lane accumulators of
     * p*log(p) by FMA (zero lanes replaced by 1 before the log), then one
     * horizontal add per entropy and a negation. */
    float inv, hy;
    mi_consts(cases, ctrls, &inv, &hy);
    const __m512 one = _mm512_set1_ps(1.0f), vinv = _mm512_set1_ps(inv);
    __m512 hx = _mm512_setzero_ps(), hall = _mm512_setzero_ps();
    for (size_t i = 0; i < size; i += 16) {
        __m512i c = _mm512_loadu_si512((const void *)(cases_ct + i));
        __m512i t = _mm512_loadu_si512((const void *)(ctrls_ct + i));
        __mmask16 mc = _mm512_test_epi32_mask(c, c);
        __mmask16 mt = _mm512_test_epi32_mask(t, t);
        __m512 pc = _mm512_mul_ps(_mm512_cvtepi32_ps(c), vinv);
        __m512 pt = _mm512_mul_ps(_mm512_cvtepi32_ps(t), vinv);
        hall = _mm512_fmadd_ps(pc, _ZGVeN16v_logf(_mm512_mask_blend_ps(mc, one, pc)), hall);
        hall = _mm512_fmadd_ps(pt, _ZGVeN16v_logf(_mm512_mask_blend_ps(mt, one, pt)), hall);
        __m512 pa = _mm512_add_ps(pc, pt);
        hx = _mm512_fmadd_ps(pa, _ZGVeN16v_logf(_mm512_mask_blend_ps((__mmask16)(mc | mt), one, pa)), hx);
    }
    float hall_f = -_mm512_reduce_add_ps(hall);
    float hx_f = -_mm512_reduce_add_ps(hx);
    return hx_f + hy - hall_f;
}
#endif

float orc_mi_f32_ref512(const uint32_t *cases_ct, const uint32_t *ctrls_ct,
                        size_t size, unsigned cases, unsigned ctrls)
{
#if defined(__x86_64__)
    if (orc_has_avx512() && size % 16 == 0)
        return mi_ref512_impl(cases_ct, ctrls_ct, size, cases, ctrls);
#endif
    return orc_mi_f32_base(cases_ct, ctrls_ct, size, cases, ctrls);
}

double orc_mi_f64(const uint32_t *cases_ct, const uint32_t *ctrls_ct,
                  size_t size, unsigned cases, unsigned ctrls)
{
    /* MI = H(X) + H(Y) - H(X,Y) in nats (MutualInformation.h:13-31), every
     * term from the exact integer counts in double precision. */
    double n = (double)cases + (double)ctrls;
    double hy = 0.0, hx = 0.0, hall = 0.0;
    if (cases) hy -= (cases / n) * log(cases / n);
    if (ctrls) hy -= (ctrls / n) * log(ctrls / n);
    for (size_t i = 0; i < size; ++i) {
        double pc = cases_ct[i] / n, pt = ctrls_ct[i] / n, pa = pc + pt;
        if (pc > 0) hall -= pc * log(pc);
        if (pt > 0) hall -= pt * log(pt);
        if (pa > 0) hx -= pa * log(pa);
    }
    return hx + hy - hall;
}

/* ------------------------------------------------------------------------ */
/* Distribution                                                              */
/* ------------------------------------------------------------------------ */

typedef struct {
    int n, k, step;
    int c[64];
} dist_it;

/* Distribution.h:60-78 (increment_left) */
static void dist_increment_left(dist_it *it)
{
    int p = it->k - 2;
    while (p >= 0 && it->c[p] + 1 > it->n - (it->k - p)) --p;
    if (p < 0) {
        it->c[0] = it->n;
    } else {
        ++it->c[p];
        for (int i = p + 1; i < it->k - 1; ++i) it->c[i] = it->c[i - 1] + 1;
    }
}

/* Distribution.h:80-87 (increment): advance x positions in lex order */
static void dist_increment(dist_it *it, int x)
{
    it->c[it->k - 1] += x;
    while (it->c[0] < it->n && it->c[it->k - 1] >= it->n) {
        dist_increment_left(it);
        if (it->k >= 2)
            it->c[it->k - 1] += (it->c[it->k - 2] + 1) - it->n;
    }
}

static void dist_begin(dist_it *it, int n, int k, int step, int offset)
{
    /* Distribution.h:101-108: start at 0,1,..,k-1 then advance offset */
    it->n = n;
    it->k = k;
    it->step = step;
    for (int i = 0; i < k; ++i) it->c[i] = i;
    dist_increment(it, offset);
}

static int dist_valid(const dist_it *it) { return it->c[0] < it->n; }

size_t orc_dist_prefixes(int n, int k1, int step, int offset, int *out,
                         size_t max)
{
    if (k1 < 1 || k1 > 64 || step < 1) return 0;
    dist_it it;
    dist_begin(&it, n, k1, step, offset);
    size_t cnt = 0;
    while (dist_valid(&it)) {
        if (out && cnt < max) memcpy(out + cnt * k1, it.c, k1 * sizeof(int));
        ++cnt;
        dist_increment(&it, step);
    }
    return cnt;
}

/* ------------------------------------------------------------------------ */
/* Search                                                                    */
/* ------------------------------------------------------------------------ */

typedef struct {
    float val;
    int c[16];
} orc_result;

/* Result.h:22: order by value only; MaxArray.h:31-46 keeps first-seen on
 * equal value (strict > to replace the minimum). */
typedef struct {
    orc_result *a;
    size_t max, size, min_pos;
} maxarray;

static void ma_add(maxarray *m, const orc_result *r)
{
    if (m->size < m->max) {
        m->a[m->size] = *r;
        if (r->val < m->a[m->min_pos].val) m->min_pos = m->size;
        ++m->size;
    } else if (m->max && r->val > m->a[m->min_pos].val) {
        m->a[m->min_pos] = *r;
        size_t mp = 0;
        for (size_t i = 1; i < m->max; ++i)
            if (m->a[i].val < m->a[mp].val) mp = i;
        m->min_pos = mp;
    }
}

typedef struct {
    const orc_dataset *d;
    int order, step, offset, mi_mode, k;
    maxarray ma;
    uint64_t evaluated;
} thread_arg;

#if defined(__x86_64__)
/* VPOPCNTDQ restatement of src/avx512vpopcntdq/gt_popcnt.cpp:11-46: 8 lanes
 * of 64-bit accumulators per cell, one horizontal sum at the end. */
__attribute__((target("avx512f,avx512vpopcntdq"))) static void
popcnt512(const uint64_t *t1, size_t size1, const uint64_t *t2, size_t words,
          uint32_t *ct)
{
    for (size_t i = 0; i < size1; ++i)
        for (size_t j = 0; j < 3; ++j) {
            __m512i acc = _mm512_setzero_si512();
            size_t w = 0;
            for (; w + 8 <= words; w += 8) {
                __m512i a = _mm512_loadu_si512((const void *)(t1 + i * words + w));
                __m512i b = _mm512_loadu_si512((const void *)(t2 + j * words + w));
                acc = _mm512_add_epi64(acc, _mm512_popcnt_epi64(_mm512_and_si512(a, b)));
            }
            uint64_t c = (uint64_t)_mm512_reduce_add_epi64(acc);
            for (; w < words; ++w)
                c += (uint64_t)__builtin_popcountll(t1[i * words + w] & t2[j * words + w]);
            ct[i * 3 + j] = (uint32_t)c;
        }
}

__attribute__((target("avx512f"))) static void
combine512(const uint64_t *t1, size_t size1, const uint64_t *t2, size_t words,
           uint64_t *out)
{
    for (size_t i = 0; i < size1; ++i)
        for (size_t j = 0; j < 3; ++j) {
            size_t w = 0;
            for (; w + 8 <= words; w += 8)
                _mm512_storeu_si512((void *)(out + (i * 3 + j) * words + w),
                                    _mm512_and_si512(_mm512_loadu_si512((const void *)(t1 + i * words + w)),
                                                     _mm512_loadu_si512((const void *)(t2 + j * words + w))));
            for (; w < words; ++w)
                out[(i * 3 + j) * words + w] = t1[i * words + w] & t2[j * words + w];
        }
}
#endif

static void do_popcnt(const uint64_t *t1, size_t size1, const uint64_t *t2,
                      size_t words, uint32_t *ct, size_t ct_size, int fast)
{
#if defined(__x86_64__)
    if (fast) {
        popcnt512(t1, size1, t2, words, ct);
        for (size_t i = size1 * 3; i < ct_size; ++i) ct[i] = 0;
        return;
    }
#endif
    (void)fast;
    orc_popcnt(t1, size1, t2, words, ct, ct_size);
}

static void do_combine(const uint64_t *t1, size_t size1, const uint64_t *t2,
                       size_t words, uint64_t *out, int fast)
{
#if defined(__x86_64__)
    if (fast) { combine512(t1, size1, t2, words, out); return; }
#endif
    (void)fast;
    orc_combine(t1, size1, t2, words, out);
}

static void *search_thread(void *p)
{
    /* ThreadedSearch.h:142-182 (nonsegmented_search) */
    thread_arg *a = (thread_arg *)p;
    const orc_dataset *d = a->d;
    int k = a->order, fast = orc_has_avx512();
    size_t cw = d->cases_words, tw = d->ctrls_words;
    size_t psize = 1;
    for (int i = 0; i < k - 1; ++i) psize *= 3;
    size_t ctsz = orc_ct_size(k, 64);
    uint64_t *bufc[2], *buft[2];
    for (int b = 0; b < 2; ++b) {
        bufc[b] = (uint64_t *)aligned_alloc(64, (psize * cw + 8) * 8);
        buft[b] = (uint64_t *)aligned_alloc(64, (psize * tw + 8) * 8);
    }
    uint32_t *cc = (uint32_t *)aligned_alloc(64, ctsz * 4 + 64);
    uint32_t *ct = (uint32_t *)aligned_alloc(64, ctsz * 4 + 64);
    dist_it it;
    dist_begin(&it, (int)d->snps, k - 1, a->step, a->offset);
    orc_result r;
    while (dist_valid(&it)) {
        const uint64_t *pc = snp_cases(d, it.c[0]), *pt = snp_ctrls(d, it.c[0]);
        size_t cur = 3;
        for (int i = 1; i < k - 1; ++i) {
            int o = i & 1;
            do_combine(pc, cur, snp_cases(d, it.c[i]), cw, bufc[o], fast);
            do_combine(pt, cur, snp_ctrls(d, it.c[i]), tw, buft[o], fast);
            pc = bufc[o];
            pt = buft[o];
            cur *= 3;
        }
        for (int c = 0; c < k - 1; ++c) r.c[c] = it.c[c];
        for (int i = it.c[k - 2] + 1; i < (int)d->snps; ++i) {
            do_popcnt(pc, cur, snp_cases(d, i), cw, cc, ctsz, fast);
            do_popcnt(pt, cur, snp_ctrls(d, i), tw, ct, ctsz, fast);
            r.c[k - 1] = i;
            if (a->mi_mode == 0)
                r.val = orc_mi_f32_ref512(cc, ct, ctsz, (unsigned)d->cases, (unsigned)d->ctrls);
            else if (a->mi_mode == 1)
                r.val = orc_mi_f32_base(cc, ct, ctsz, (unsigned)d->cases, (unsigned)d->ctrls);
            else
                r.val = (float)orc_mi_f64(cc, ct, ctsz, (unsigned)d->cases, (unsigned)d->ctrls);
            ma_add(&a->ma, &r);
            ++a->evaluated;
        }
        dist_increment(&it, a->step);
    }
    for (int b = 0; b < 2; ++b) { free(bufc[b]); free(buft[b]); }
    free(cc);
    free(ct);
    return NULL;
}

static int g_k_for_sort;
static int cmp_result(const void *x, const void *y)
{
    /* value descending, then combination lexicographic ascending */
    const orc_result *a = (const orc_result *)x, *b = (const orc_result *)y;
    if (a->val > b->val) return -1;
    if (a->val < b->val) return 1;
    for (int i = 0; i < g_k_for_sort; ++i) {
        if (a->c[i] < b->c[i]) return -1;
        if (a->c[i] > b->c[i]) return 1;
    }
    return 0;
}

long orc_search(const orc_dataset *d, int order, int step, int offset,
                unsigned outputs, int threads, int mi_mode, int *combos,
                float *mi, uint64_t *evaluated)
{
    /* ThreadedSearch.h:207-246: thread i scans distribution.layer(T, i);
     * the per-thread lists are concatenated, sorted and truncated. */
    if (order < 2 || order > 16 || threads < 1 || step < 1) return -1;
    thread_arg *args = (thread_arg *)calloc((size_t)threads, sizeof(thread_arg));
    pthread_t *tid = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
    for (int t = 0; t < threads; ++t) {
        args[t].d = d;
        args[t].order = order;
        args[t].k = order;
        args[t].step = step * threads;          /* Distribution.h:172-176 */
        args[t].offset = offset * threads + t;
        args[t].mi_mode = mi_mode;
        args[t].ma.max = outputs;
        args[t].ma.a = (orc_result *)calloc(outputs + 1, sizeof(orc_result));
        pthread_create(&tid[t], NULL, search_thread, &args[t]);
    }
    size_t total = 0;
    uint64_t ev = 0;
    for (int t = 0; t < threads; ++t) {
        pthread_join(tid[t], NULL);
        total += args[t].ma.size;
        ev += args[t].evaluated;
    }
    orc_result *all = (orc_result *)calloc(total + 1, sizeof(orc_result));
    size_t pos = 0;
    for (int t = 0; t < threads; ++t) {
        memcpy(all + pos, args[t].ma.a, args[t].ma.size * sizeof(orc_result));
        pos += args[t].ma.size;
        free(args[t].ma.a);
    }
    g_k_for_sort = order;
    qsort(all, total, sizeof(orc_result), cmp_result);
    size_t n = total < outputs ? total : outputs;
    for (size_t i = 0; i < n; ++i) {
        if (combos) memcpy(combos + i * order, all[i].c, order * sizeof(int));
        if (mi) mi[i] = all[i].val;
    }
    if (evaluated) *evaluated = ev;
    free(all);
    free(args);
    free(tid);
    return (long)n;
}

int orc_has_avx512(void)
{
#if defined(__x86_64__)
    __builtin_cpu_init();
    return __builtin_cpu_supports("avx512f") &&
           __builtin_cpu_supports("avx512vpopcntdq") &&
           __builtin_cpu_supports("avx512bw") &&
           __builtin_cpu_supports("avx512dq");
#else
    return 0;
#endif
}

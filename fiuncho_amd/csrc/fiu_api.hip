// C ABI of fiuncho_amd (declared in include/fiuncho_amd.h): dataset upload,
// the chunked search driver and the batch operators.
#include "../../include/fiuncho_amd.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "fiu_internal.h"
#include "fiu_kernels.h"
#include "fiu_loader.h"
#include "fiu_synth.h"

using namespace fiu;

namespace fiu {

std::string &error_slot()
{
    thread_local std::string s;
    return s;
}

int &current_device()
{
    thread_local int d = 0;
    return d;
}

uint64_t binom_sat(uint64_t n, uint64_t k)
{
    if (k > n) return 0;
    k = std::min(k, n - k);
    unsigned __int128 r = 1;
    for (uint64_t i = 1; i <= k; ++i) {
        r = r * (n - k + i) / i;
        if (r > (unsigned __int128)UINT64_MAX) return UINT64_MAX;
    }
    return (uint64_t)r;
}

// Lexicographic unranking of an m-combination of [0, n) on the host (the
// device's unrank, with binom_sat for the table).
void unrank_host(uint64_t r, uint64_t n, int m, int32_t *out)
{
    uint64_t x = binom_sat(n, (uint64_t)m) - 1 - r;
    int64_t hi = (int64_t)n - 1;
    for (int i = 0; i < m; ++i) {
        const uint64_t j = (uint64_t)(m - i);
        int64_t lo = 0, h = hi;   // largest lo <= hi with C(lo, j) <= x
        while (lo < h) {
            const int64_t mid = (lo + h + 1) / 2;
            if (binom_sat((uint64_t)mid, j) <= x) lo = mid;
            else h = mid - 1;
        }
        x -= binom_sat((uint64_t)lo, j);
        out[i] = (int32_t)((int64_t)n - 1 - lo);
        hi = lo - 1;
    }
}

}  // namespace fiu

fiu_dataset::~fiu_dataset()
{
    if (readback) (void)hipHostFree(readback);
    for (auto e : events) (void)hipEventDestroy(e);
    if (stream && own_stream) (void)hipStreamDestroy(stream);
    for (auto &r : retired) {
        if (r.second) r.second->release(r.first);
        else (void)hipFree(r.first);
    }
}

namespace fiu {

namespace {

struct Shape {
    int n, n_pad, q0, q1;
    uint64_t off1, total;   // uint4 units
};

Shape shape_of(size_t snps, size_t cw, size_t tw)
{
    Shape s;
    s.n = (int)snps;
    s.n_pad = (s.n + 63) / 64 * 64;
    s.q0 = (int)((cw + 1) / 2);
    s.q1 = (int)((tw + 1) / 2);
    s.off1 = (uint64_t)s.q0 * 3 * s.n_pad;
    s.total = s.off1 + (uint64_t)s.q1 * 3 * s.n_pad;
    return s;
}

void bind_stream(fiu_dataset *d, hipStream_t stream)
{
    if (d->snps > (size_t)INT32_MAX - 64) fail(FIU_ERR_UNSUPPORTED, "more SNPs than INT_MAX");
    check(hipSetDevice(d->device), "hipSetDevice");
    if (stream) {
        d->stream = stream;
        d->own_stream = false;
    } else {
        check(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking), "hipStreamCreate");
        d->own_stream = true;
    }
}

// Device image from reference-layout words already on the device: the
// transposes into the SNP-minor quads, then row counts and coverage flags.
// Returns once the device no longer reads `words`.
void finish_upload(fiu_dataset *d, const uint64_t *words)
{
    const Shape sh = shape_of(d->snps, d->cw, d->tw);
    const int n = sh.n, n_pad = sh.n_pad;
    d->bits.alloc(sh.total);
    check(launch_transpose(words, d->bits.p, n, n_pad, 0, d->cw, d->tw, sh.q0, 0, d->stream), "transpose");
    check(launch_transpose(words, d->bits.p, n, n_pad, 1, d->cw, d->tw, sh.q1, sh.off1, d->stream), "transpose");
    // the caller releases (and may reuse) `words` once the transposes read it
    check(hipStreamSynchronize(d->stream), "transpose sync");
    DevData &D = d->dev;
    D.bits = d->bits.p;
    D.class_off[0] = 0;
    D.class_off[1] = sh.off1;
    D.quads[0] = sh.q0;
    D.quads[1] = sh.q1;
    D.n = n;
    D.n_pad = n_pad;
    D.cases = d->cases;
    D.ctrls = d->ctrls;
    d->singles.alloc((size_t)n_pad * 6);
    d->flags.alloc(4);
    uint32_t zero[4] = {0, 0, 0, 0};
    check(hipMemcpyAsync(d->flags.p, zero, sizeof(zero), hipMemcpyHostToDevice, d->stream), "flags");
    D.singles = d->singles.p;
    if (n > 0) check(launch_singles(d->bits.p, D, d->singles.p, d->flags.p, d->stream), "singles");
    uint32_t fl[4];
    check(hipMemcpyAsync(fl, d->flags.p, sizeof(fl), hipMemcpyDeviceToHost, d->stream), "flags");
    check(hipStreamSynchronize(d->stream), "singles sync");
    D.uniform = (n > 0 && fl[0] == 0) ? 1 : 0;
    D.total[0] = fl[1];
    D.total[1] = fl[2];
}

void upload(fiu_dataset *d, const uint64_t *words, hipStream_t stream)
{
    bind_stream(d, stream);
    const size_t nwords = d->snps * 3 * (d->cw + d->tw);
    DBuf<uint64_t> staging;
    staging.arena = d->arena;
    staging.alloc(nwords);
    if (nwords) check(hipMemcpyAsync(staging.p, words, nwords * 8, hipMemcpyHostToDevice, d->stream), "upload");
    finish_upload(d, staging.p);
}

}  // namespace

size_t dataset_bytes(size_t snps, size_t cw, size_t tw)
{
    const Shape sh = shape_of(snps, cw, tw);
    auto up = [](size_t b) { return (b + Arena::kAlign - 1) / Arena::kAlign * Arena::kAlign; };
    return up(std::max<uint64_t>(sh.total, 1) * 16) + up((size_t)sh.n_pad * 24) + up(16);
}

size_t staging_bytes(size_t snps, size_t cw, size_t tw)
{
    return (std::max<size_t>(snps * 3 * (cw + tw), 1) * 8 + Arena::kAlign - 1) / Arena::kAlign * Arena::kAlign;
}

fiu_dataset *create_dataset(const uint64_t *words, size_t snps, size_t cw, size_t tw, uint32_t cases,
                            uint32_t ctrls, Arena *arena, hipStream_t stream, bool keep_host_words)
{
    if (!words && snps) fail(FIU_ERR_INVALID, "words is NULL");
    if (snps && (cw == 0 || tw == 0) && (cases || ctrls)) fail(FIU_ERR_INVALID, "zero words per row");
    if ((uint64_t)cases > cw * 64 || (uint64_t)ctrls > tw * 64)
        fail(FIU_ERR_INVALID, "more samples than bits per row");
    auto d = std::make_unique<fiu_dataset>();
    d->device = current_device();
    d->use_arena(arena);
    d->snps = snps;
    d->cw = cw;
    d->tw = tw;
    d->cases = cases;
    d->ctrls = ctrls;
    if (keep_host_words) {
        d->host_words.assign(words, words + snps * 3 * (cw + tw));
        upload(d.get(), d->host_words.data(), stream);
    } else {
        upload(d.get(), words, stream);
    }
    return d.release();
}

fiu_dataset *dataset_from_genotypes(const int8_t *geno, size_t snps, size_t samples, const uint8_t *is_ctrl,
                                    bool reader_flush, Arena *arena, hipStream_t stream)
{
    if (!samples) fail(FIU_ERR_INVALID, "no samples");
    if (!is_ctrl) fail(FIU_ERR_INVALID, "is_ctrl is NULL");
    if (!geno && snps) fail(FIU_ERR_INVALID, "geno is NULL");
    if (samples > (size_t)INT32_MAX) fail(FIU_ERR_UNSUPPORTED, "more samples than INT_MAX");
    // per-class sample lists (file order within a class), cases first; not
    // needed when each class is one run of samples
    std::vector<int32_t> idx(samples);
    size_t ctrls = 0, changes = 0;
    for (size_t j = 0; j < samples; ++j) {
        ctrls += is_ctrl[j] != 0;
        changes += j && (is_ctrl[j] != 0) != (is_ctrl[j - 1] != 0);
    }
    const size_t cases = samples - ctrls;
    {
        size_t a = 0, b = cases;
        for (size_t j = 0; j < samples; ++j) idx[is_ctrl[j] ? b++ : a++] = (int32_t)j;
    }
    const bool runs = changes <= 1;
    const uint32_t n_class[2] = {(uint32_t)cases, (uint32_t)ctrls};
    const uint32_t start[2] = {runs ? (uint32_t)idx[0] : 0u, runs ? (uint32_t)idx[cases] : 0u};
    const size_t cw = packed_words(cases), tw = packed_words(ctrls);
    if (!cases || !ctrls) fail(FIU_ERR_INVALID, "zero words per row: one class has no samples");
    if (!snps) return create_dataset(nullptr, 0, cw, tw, (uint32_t)cases, (uint32_t)ctrls, arena, stream);
    std::vector<uint8_t> zero;
    if (reader_flush) {
        for (const auto &q : flush_quirk_words(is_ctrl, samples)) {
            if (zero.empty()) zero.assign(cw + tw, 0);
            zero[q.first ? cw + q.second : q.second] = 1;
        }
    }
    auto d = std::make_unique<fiu_dataset>();
    d->device = current_device();
    d->use_arena(arena);
    d->snps = snps;
    d->cw = cw;
    d->tw = tw;
    d->cases = (uint32_t)cases;
    d->ctrls = (uint32_t)ctrls;
    bind_stream(d.get(), stream);
    const size_t nwords = snps * 3 * (cw + tw);
    DBuf<uint64_t> words;
    words.arena = d->arena;
    words.alloc(nwords);
    {
        DBuf<int8_t> g;
        DBuf<int32_t> di;
        DBuf<uint8_t> dz;
        g.arena = di.arena = dz.arena = d->arena;
        g.alloc((snps * samples + 15) / 16 * 16);   // the kernel reads whole 16-byte blocks
        check(hipMemcpyAsync(g.p, geno, snps * samples, hipMemcpyHostToDevice, d->stream), "genotype upload");
        if (!runs) {
            di.alloc(samples);
            check(hipMemcpyAsync(di.p, idx.data(), samples * 4, hipMemcpyHostToDevice, d->stream), "sample lists");
        }
        if (!zero.empty()) {
            dz.alloc(zero.size());
            check(hipMemcpyAsync(dz.p, zero.data(), zero.size(), hipMemcpyHostToDevice, d->stream), "flush words");
        }
        check(launch_pack_words(g.p, snps, samples, runs ? nullptr : di.p, n_class, start, cw, tw,
                                zero.empty() ? nullptr : dz.p, words.p, d->stream),
              "pack");
        d->host_words.resize(nwords);
        check(hipMemcpyAsync(d->host_words.data(), words.p, nwords * 8, hipMemcpyDeviceToHost, d->stream),
              "words download");
        check(hipStreamSynchronize(d->stream), "pack sync");
    }
    finish_upload(d.get(), words.p);
    return d.release();
}

namespace {

size_t align_up(size_t b) { return (b + Arena::kAlign - 1) / Arena::kAlign * Arena::kAlign; }

// Row popcounts, disjointness of each SNP's rows and equality of their union
// with SNP 0's union, per class -- what k_singles computes on the device.
void host_singles(const uint64_t *words, size_t snps, size_t cw, size_t tw, uint32_t *singles, bool &uniform,
                  uint32_t total[2])
{
    const size_t per = 3 * (cw + tw);
    bool bad = false;
    for (int cls = 0; cls < 2; ++cls) {
        const size_t W = cls ? tw : cw, off = cls ? 3 * cw : 0;
        std::vector<uint64_t> ref(W, 0);
        if (snps) {
            const uint64_t *r = words + off;
            for (size_t w = 0; w < W; ++w) ref[w] = r[w] | r[W + w] | r[2 * W + w];
        }
        uint32_t tot = 0;
        for (size_t w = 0; w < W; ++w) tot += (uint32_t)__builtin_popcountll(ref[w]);
        total[cls] = snps ? tot : 0;
        for (size_t s = 0; s < snps; ++s) {
            const uint64_t *r0 = words + s * per + off, *r1 = r0 + W, *r2 = r1 + W;
            uint32_t c0 = 0, c1 = 0, c2 = 0;
            uint64_t b = 0;
            for (size_t w = 0; w < W; ++w) {
                const uint64_t a = r0[w], bb = r1[w], c = r2[w];
                c0 += (uint32_t)__builtin_popcountll(a);
                c1 += (uint32_t)__builtin_popcountll(bb);
                c2 += (uint32_t)__builtin_popcountll(c);
                b |= (a & bb) | (a & c) | (bb & c) | ((a | bb | c) ^ ref[w]);
            }
            singles[(s * 2 + cls) * 3 + 0] = c0;
            singles[(s * 2 + cls) * 3 + 1] = c1;
            singles[(s * 2 + cls) * 3 + 2] = c2;
            bad |= b != 0;
        }
    }
    uniform = snps > 0 && !bad;
}

}  // namespace

size_t staged_host_bytes(size_t snps, size_t cw, size_t tw, int order)
{
    const Shape sh = shape_of(snps, cw, tw);
    const size_t M = (size_t)std::max(order, 2);   // binomial rows up to k (prefix ranks and keys)
    return align_up(snps * 3 * (cw + tw) * 8) + align_up((size_t)sh.n_pad * 24) + align_up((M + 1) * (snps + 1) * 8);
}

fiu_dataset *stage_dataset(const uint64_t *words, size_t snps, size_t cw, size_t tw, uint32_t cases,
                           uint32_t ctrls, int order, bool want_fp4, Arena *arena, Arena *pinned, StagedUpload &up)
{
    if (!words && snps) fail(FIU_ERR_INVALID, "words is NULL");
    if (snps && (cw == 0 || tw == 0) && (cases || ctrls)) fail(FIU_ERR_INVALID, "zero words per row");
    if ((uint64_t)cases > cw * 64 || (uint64_t)ctrls > tw * 64)
        fail(FIU_ERR_INVALID, "more samples than bits per row");
    if (snps > (size_t)INT32_MAX - 64) fail(FIU_ERR_UNSUPPORTED, "more SNPs than INT_MAX");
    const Shape sh = shape_of(snps, cw, tw);
    const int M = std::max(order, 2);   // binomial rows up to k (prefix ranks and keys)
    up = StagedUpload();
    up.n_words = snps * 3 * (cw + tw);
    up.off_singles = align_up(up.n_words * 8);
    up.off_binom = up.off_singles + align_up((size_t)sh.n_pad * 24);
    up.binom_n = (size_t)(M + 1) * (snps + 1);
    up.host_bytes = up.off_binom + align_up(up.binom_n * 8);
    up.host = (char *)pinned->alloc(up.host_bytes);
    if (!up.host) return nullptr;
    auto d = std::make_unique<fiu_dataset>();
    d->device = current_device();
    d->use_arena(arena);
    d->snps = snps;
    d->cw = cw;
    d->tw = tw;
    d->cases = cases;
    d->ctrls = ctrls;
    d->own_stream = false;
    try {
        d->bits.alloc(sh.total);
        d->singles.alloc((size_t)sh.n_pad * 6);
        d->binom.alloc(up.binom_n);
        up.staging_d = (uint64_t *)arena->alloc(std::max<size_t>(up.n_words, 1) * 8);
        if (!up.staging_d) fail(FIU_ERR_NOMEM, "no room for the upload staging");
    } catch (const Failure &) {
        pinned->release(up.host);
        up.host = nullptr;
        return nullptr;  // the dataset's buffers go back to the arena with it
    }
    // host side of the preparation
    std::memcpy(up.host, words, up.n_words * 8);
    uint32_t *sg = (uint32_t *)(up.host + up.off_singles);
    std::memset(sg, 0, (size_t)sh.n_pad * 24);
    bool uniform = false;
    uint32_t total[2] = {0, 0};
    host_singles(words, snps, cw, tw, sg, uniform, total);
    uint64_t *bn = (uint64_t *)(up.host + up.off_binom);
    for (int j = 0; j <= M; ++j)
        for (uint64_t x = 0; x <= snps; ++x) bn[j * (snps + 1) + x] = binom_sat(x, (uint64_t)j);
    d->binom_m = M;
    DevData &D = d->dev;
    D.bits = d->bits.p;
    D.class_off[0] = 0;
    D.class_off[1] = sh.off1;
    D.quads[0] = sh.q0;
    D.quads[1] = sh.q1;
    D.n = sh.n;
    D.n_pad = sh.n_pad;
    D.cases = cases;
    D.ctrls = ctrls;
    D.singles = d->singles.p;
    D.uniform = uniform ? 1 : 0;
    D.total[0] = total[0];
    D.total[1] = total[1];
    if (order == 3 && uniform) {
        d->pair.arena = arena;
        if (pair_table_bytes(snps) <= arena->largest_free()) {
            try {
                d->pair.alloc(snps * (snps - 1) / 2 * 8);
                up.pair = true;
                d->pair_ready = true;   // in stream order, before any scan of it
            } catch (const Failure &) {
                up.pair = false;        // no room: the scans fall back to the direct kind
            }
        }
    }
    if (want_fp4 && mfma_dataset(d.get()) && (order == 2 || (order == 3 && up.pair))) {
        const size_t units = fp4_image_units(D, 0) + fp4_image_units(D, 1);
        if (units * 16 <= arena->largest_free()) {
            try {
                d->fp4.alloc(units);
                up.fp4 = true;
                d->fp4_ready = true;    // in stream order, before any scan of it
            } catch (const Failure &) {
                up.fp4 = false;         // no room: the popcount kinds scan it
            }
        }
    }
    return d.release();
}

void enqueue_upload(fiu_dataset *d, StagedUpload &up, hipStream_t s)
{
    const Shape sh = shape_of(d->snps, d->cw, d->tw);
    // every copy below stays inside its pinned block or device buffer
    check_range(up.host + up.off_singles, (size_t)sh.n_pad * 24, up.host, up.host_bytes, "row-count upload");
    check_range(up.host + up.off_binom, up.binom_n * 8, up.host, up.host_bytes, "binomial upload");
    check_range(d->singles.p, (size_t)sh.n_pad * 24, d->singles.p, d->singles.n * 4, "row-count upload");
    check_range(d->binom.p, up.binom_n * 8, d->binom.p, d->binom.n * 8, "binomial upload");
    if (up.n_words) {
        check_range(up.host, up.n_words * 8, up.host, up.host_bytes, "table upload");
        check(hipMemcpyAsync(up.staging_d, up.host, up.n_words * 8, hipMemcpyHostToDevice, s), "upload");
    }
    check(launch_transpose(up.staging_d, d->bits.p, sh.n, sh.n_pad, 0, d->cw, d->tw, sh.q0, 0, s), "transpose");
    check(launch_transpose(up.staging_d, d->bits.p, sh.n, sh.n_pad, 1, d->cw, d->tw, sh.q1, sh.off1, s), "transpose");
    check(hipMemcpyAsync(d->singles.p, up.host + up.off_singles, (size_t)sh.n_pad * 24, hipMemcpyHostToDevice, s),
          "row counts");
    check(hipMemcpyAsync(d->binom.p, up.host + up.off_binom, up.binom_n * 8, hipMemcpyHostToDevice, s), "binomials");
    if (up.pair) {
        ScanParams P{};
        P.j0 = 0;
        P.j1 = (int64_t)d->snps;
        P.step = 1;
        P.offset = 0;
        P.total_prefixes = d->snps;
        P.binom = d->binom.p;
        check(launch_pair_table(d->dev, P, d->pair.p, s), "pair table launch");
    }
    if (up.fp4) check(launch_fp4_image(d->dev, d->fp4.p, s), "fp4 image launch");
}

namespace {
std::mutex g_tuning_mu;
std::vector<fiu_tune_entry> g_tuning;
}  // namespace

bool valid_variant(int algo, int order, int variant)
{
    if (variant == 0) return true;
    switch (algo) {
    case FIU_ALGO_MARGINAL:
        return order <= 3 && (variant == 2 || variant == kMarginalPW || (order == 2 && variant == 8));
    case FIU_ALGO_MFMA: return order == 3 ? (variant >= 3 && variant <= 6) : (variant >= 1 && variant <= 4);
    default: return false;
    }
}

bool tuned(int order, uint64_t snps, uint64_t samples, fiu_tune_entry &out)
{
    std::lock_guard<std::mutex> lk(g_tuning_mu);
    for (const auto &e : g_tuning)
        if (e.order == order && snps >= e.snps_lo && snps < e.snps_hi && samples >= e.samples_lo &&
            samples < e.samples_hi) {
            out = e;
            return true;
        }
    return false;
}

void mi_constants(uint32_t cases, uint32_t ctrls, bool complete, ScanParams &P)
{
    P.bounded = complete ? 1u : 0u;
    const double n = (double)cases + (double)ctrls;
    double hy = 0.0;
    if (cases) hy -= ((double)cases / n) * std::log((double)cases / n);
    if (ctrls) hy -= ((double)ctrls / n) * std::log((double)ctrls / n);
    P.n64 = n;
    P.hy64 = hy;
    P.hy = (float)hy;
    P.inv_n = (float)(1.0 / n);
    P.frac_c = (float)((double)cases / n);
    P.chi_scale = cases && ctrls ? (float)(n / ((double)cases * (double)ctrls)) : 0.0f;
    // bound on |screen - exact| (a few 1e-6 nats) with a 5x margin
    P.delta = 2e-5f;
}

const uint64_t *ensure_binom(fiu_dataset *d, int m, hipStream_t s)
{
    std::lock_guard<std::mutex> lk(d->prep_mu);
    if (d->binom_m >= m) return d->binom.p;
    const uint64_t n = d->snps, stride = n + 1;
    std::vector<uint64_t> h((uint64_t)(m + 1) * stride);
    for (int j = 0; j <= m; ++j)
        for (uint64_t x = 0; x <= n; ++x) h[j * stride + x] = binom_sat(x, (uint64_t)j);
    // scans queued with the smaller table (another thread's, or the scan
    // server's) may still read it: it is kept until the dataset goes
    if (d->binom.p) d->retired.emplace_back((void *)d->binom.detach(), d->binom.arena);
    d->binom.alloc(h.size());
    check(hipMemcpyAsync(d->binom.p, h.data(), h.size() * 8, hipMemcpyHostToDevice, s), "binom upload");
    check(hipStreamSynchronize(s), "binom sync");
    d->binom_m = m;
    return d->binom.p;
}

bool ensure_pair_table(fiu_dataset *d, hipStream_t s)
{
    const uint64_t *binom = ensure_binom(d, 1, s);
    std::lock_guard<std::mutex> lk(d->prep_mu);
    if (d->pair_ready) return true;
    const uint64_t n = d->snps, pairs = n * (n - 1) / 2;
    if (d->arena) {
        if (pairs * 32 > d->arena->largest_free()) return false;
    } else {
        size_t free_b = 0, total_b = 0;
        check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        if (pairs * 32 + (64ull << 20) > free_b) return false;
    }
    d->pair.alloc(pairs * 8);
    ScanParams P{};
    P.j0 = 0;
    P.j1 = (int64_t)n;
    P.step = 1;
    P.offset = 0;
    P.total_prefixes = n;
    P.binom = binom;
    check(launch_pair_table(d->dev, P, d->pair.p, s), "pair table launch");
    check(hipStreamSynchronize(s), "pair table sync");
    d->pair_ready = true;
    return true;
}

bool ensure_fp4_image(fiu_dataset *d, hipStream_t s)
{
    std::lock_guard<std::mutex> lk(d->prep_mu);
    if (d->fp4_ready) return true;
    const uint64_t units = fp4_image_units(d->dev, 0) + fp4_image_units(d->dev, 1);
    if (d->arena) {
        if (units * 16 > d->arena->largest_free()) return false;
    } else {
        size_t free_b = 0, total_b = 0;
        check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        if (units * 16 + (64ull << 20) > free_b) return false;
    }
    d->fp4.alloc(units);
    check(launch_fp4_image(d->dev, d->fp4.p, s), "fp4 image launch");
    check(hipStreamSynchronize(s), "fp4 image sync");
    d->fp4_ready = true;
    return true;
}

// The prefixes that share their first m-1 indices form a run of consecutive
// ranks whose last index increases by one; each run is summed in closed form.
// The (m-1)-combinations heading the runs are walked in lexicographic order
// (O(C(n, m-1)) steps, no allocation).
uint64_t count_combinations(uint64_t n, int m, uint64_t step, uint64_t offset)
{
    constexpr int kMaxM = 32;
    if (m < 1 || m > kMaxM) return 0;
    // far fewer prefixes in the share than run heads (a large step): unrank
    // each prefix of the share instead
    const uint64_t all = binom_sat(n, (uint64_t)m);
    const uint64_t n_pref = all > offset ? (all - offset + step - 1) / step : 0;
    if (n_pref < binom_sat(n, (uint64_t)m - 1) / 64) {
        int32_t c[kMaxM];
        uint64_t sum = 0;
        for (uint64_t j = 0; j < n_pref; ++j) {
            unrank_host(offset + j * step, n, m, c);
            sum += n - 1 - (uint64_t)c[m - 1];
        }
        return sum;
    }
    uint64_t total = 0, rank = 0;
    int64_t c[kMaxM];
    const int h = m - 1;   // indices that head a run
    for (int i = 0; i < h; ++i) c[i] = i;
    if (h > 0 && (uint64_t)h > n) return 0;
    for (;;) {
        const int64_t first_last = h ? c[h - 1] + 1 : 0;
        if (first_last < (int64_t)n) {
            const uint64_t L = n - (uint64_t)first_last;   // runs of last = first_last .. n-1
            // first local index whose rank is offset + j*step, j >= 0
            const uint64_t first = rank >= offset ? (step - (rank - offset) % step) % step : offset - rank;
            if (first < L) {
                const uint64_t cnt = (L - 1 - first) / step + 1, t_last = first + (cnt - 1) * step;
                // combos for local index t: n - 1 - (first_last + t) = L - 1 - t
                total += cnt * (L - 1) - (first + t_last) * cnt / 2;
            }
            rank += L;
        }
        // next h-combination of [0, n) in lexicographic order
        int i = h - 1;
        while (i >= 0 && c[i] == (int64_t)n - h + i) --i;
        if (i < 0) break;
        ++c[i];
        for (int j = i + 1; j < h; ++j) c[j] = c[j - 1] + 1;
    }
    return total;
}

}  // namespace fiu

namespace {

// Number of prefixes at the end of the distribution whose combinations add up
// to `target`: those prefixes have the fewest last-SNP candidates, so a
// pass over them spreads over many short wavefronts.  Walks the prefixes in
// reverse lexicographic order (at most 2^26 steps).
int64_t tail_prefixes(int n, int M, uint64_t step, uint64_t offset, uint64_t total, int64_t n_pref, uint64_t target)
{
    constexpr int kMaxM = 31;
    if (n_pref <= 0 || total == 0 || M < 1 || M > kMaxM) return 0;
    if (step > 64) {
        // a large step: unrank the share's last prefixes directly
        int32_t c[kMaxM];
        uint64_t acc = 0;
        int64_t m = 0;
        for (int64_t j = n_pref - 1; j >= 0 && m < (1 << 20); --j) {
            unrank_host(offset + (uint64_t)j * step, (uint64_t)n, M, c);
            ++m;
            acc += (uint64_t)(n - 1 - c[M - 1]);
            if (acc >= target) break;
        }
        return m;
    }
    int c[kMaxM];
    for (int i = 0; i < M; ++i) c[i] = n - M + i;   // the last prefix
    uint64_t rank = total - 1, acc = 0;
    int64_t m = 0;
    for (uint64_t visited = 0; visited < (1ull << 26); ++visited) {
        if (rank >= offset && (rank - offset) % step == 0) {
            ++m;
            acc += (uint64_t)(n - 1 - c[M - 1]);
            if (acc >= target || m >= n_pref) break;
        }
        if (rank == 0) break;
        // previous M-combination: lower the rightmost index that can go
        // down, then push every later index to its maximum
        int i = M - 1;
        while (i >= 0 && c[i] == (i ? c[i - 1] + 1 : 0)) --i;
        if (i < 0) break;
        --c[i];
        for (int j = i + 1; j < M; ++j) c[j] = n - M + j;
        --rank;
    }
    return m;
}

hipStream_t stream_of(const fiu_dataset *d, const fiu_config *cfg)
{
    return cfg && cfg->stream ? (hipStream_t)cfg->stream : d->stream;
}

struct Chunk {
    int64_t j0, j1;
};

// The launches of one search, computed on demand (no plan is stored):
//   uniform:   [ci len, (ci + 1) len) of `units` (the distribution's
//              prefixes, or the order-2 matrix-core super-blocks);
//   first SNP: order 3 on the matrix cores, first SNPs [a(ci), a(ci + 1)),
//              a(i) = the first a whose first SNPs 0 .. a - 1 hold at least
//              i x target triples, so that each chunk holds about `target`
//              triples (a chunk can be empty: one first SNP above target).
struct ChunkPlan {
    bool first_snp = false;
    int64_t units = 0, len = 1;
    int64_t n = 0;
    unsigned __int128 target = 1, triples = 0;
    uint64_t count = 0;

    static unsigned __int128 c3(int64_t m) { return m < 3 ? 0 : (unsigned __int128)m * (m - 1) * (m - 2) / 6; }
    void set_uniform(int64_t u, int64_t l)
    {
        first_snp = false;
        units = u;
        len = std::max<int64_t>(1, l);
        count = u > 0 ? (uint64_t)((u + len - 1) / len) : 0;
    }
    void set_first_snp(int64_t snps, uint64_t t)
    {
        first_snp = true;
        n = snps;
        triples = c3(n);
        target = std::max<uint64_t>(t, 1);
        count = n >= 3 ? (uint64_t)((triples + target - 1) / target) : 0;
    }
    int64_t boundary(uint64_t i) const
    {
        if (i >= count) return n - 2;
        const unsigned __int128 want = (unsigned __int128)i * target;
        int64_t lo = 0, hi = n - 2;   // triples with first SNP below a: C(n, 3) - C(n - a, 3)
        while (lo < hi) {
            const int64_t mid = lo + (hi - lo) / 2;
            if (triples - c3(n - mid) >= want) hi = mid;
            else lo = mid + 1;
        }
        return lo;
    }
    Chunk at(uint64_t ci) const
    {
        if (first_snp) return {boundary(ci), boundary(ci + 1)};
        const int64_t j0 = (int64_t)ci * len;
        return {j0, std::min(units, j0 + len)};
    }
};

// Device memory one search works in, and what it may build.
class Workspace {
  public:
    virtual ~Workspace() = default;
    // binomial table C(x, j), x <= snps, j <= m
    virtual const uint64_t *binom(fiu_dataset *d, int m, hipStream_t s) = 0;
    // the dataset's pair table / fp4 image, built when missing if allowed
    virtual bool pair_table(fiu_dataset *d, hipStream_t s) = 0;
    virtual bool fp4_image(fiu_dataset *d, hipStream_t s) = 0;
    virtual uint32_t max_candidates() const = 0;
    virtual uint64_t max_chunks() const = 0;
    // whether kernels that keep counts in per-lane private memory (the 64 x
    // 64 order-3 and 64 x 128 order-2 tiles) may run: the runtime backs that
    // memory at their launch
    virtual bool private_memory() const { return true; }
    // candidate buffer (cap entries), top list (max(outputs, kMaxOutputs)),
    // state, redo list (chunks entries), histogram (kHistBins)
    virtual void buffers(fiu_dataset *d, uint32_t cap, uint64_t chunks, uint32_t outputs, Cand *&cand, Cand *&top,
                         TopState *&state, uint32_t *&redo, uint32_t *&hist) = 0;
    // outputs above kMaxOutputs: the merge's sort buffer of `entries`
    virtual Cand *merge_work(fiu_dataset *d, size_t entries) = 0;
    // orders 10-12: the count slots of the persistent all-cells kernel
    virtual uint32_t *big_cells(fiu_dataset *d, int order) = 0;
    // timing event i, or null (the search is not timed)
    virtual hipEvent_t event(fiu_dataset *d, size_t i) = 0;
    // the search state, after a synchronisation of s; where it can, the
    // workspace reads the top list's first `outputs` entries in the same
    // synchronisation, for results()
    virtual void read_state(fiu_dataset *d, const TopState *state, const Cand *top, uint32_t outputs, TopState &hs,
                            hipStream_t s)
    {
        check(hipMemcpyAsync(&hs, state, sizeof(hs), hipMemcpyDeviceToHost, s), "state read");
        check(hipStreamSynchronize(s), "search sync");
    }
    // the final top list to the caller's arrays
    virtual void results(const Cand *top, uint32_t count, int order, int n, const uint64_t *binom, int32_t *out_combos,
                         float *out_mi, hipStream_t s) = 0;
};

// fiu_search: the dataset's own buffers, grown on demand and kept for later
// calls; the pair table and the fp4 image are built when missing.
class DatasetWorkspace final : public Workspace {
  public:
    const uint64_t *binom(fiu_dataset *d, int m, hipStream_t s) override { return ensure_binom(d, m, s); }
    bool pair_table(fiu_dataset *d, hipStream_t s) override { return ensure_pair_table(d, s); }
    bool fp4_image(fiu_dataset *d, hipStream_t s) override { return ensure_fp4_image(d, s); }
    uint32_t max_candidates() const override { return UINT32_MAX; }
    uint64_t max_chunks() const override { return UINT64_MAX; }
    void buffers(fiu_dataset *d, uint32_t cap, uint64_t chunks, uint32_t outputs, Cand *&cand, Cand *&top,
                 TopState *&state, uint32_t *&redo, uint32_t *&hist) override
    {
        d->cand.alloc(cap);
        d->top.alloc(std::max<uint32_t>(outputs, kMaxOutputs));
        d->state.alloc(1);
        d->redo.alloc(std::max<uint64_t>(chunks, 1));
        d->hist.alloc(kHistBins);
        cand = d->cand.p, top = d->top.p, state = d->state.p, redo = d->redo.p, hist = d->hist.p;
    }
    uint32_t *big_cells(fiu_dataset *d, int order) override
    {
        d->big.alloc(big_cells_bytes(order) / sizeof(uint32_t));
        return d->big.p;
    }
    Cand *merge_work(fiu_dataset *d, size_t entries) override
    {
        d->mwork.alloc(entries);
        return d->mwork.p;
    }
    hipEvent_t event(fiu_dataset *d, size_t i) override
    {
        while (d->events.size() <= i) {
            hipEvent_t e;
            check(hipEventCreate(&e), "hipEventCreate");
            d->events.push_back(e);
        }
        return d->events[i];
    }
    // up to kMaxOutputs results: the state and the list in one copy each
    // into the dataset's pinned readback block, one synchronisation
    void read_state(fiu_dataset *d, const TopState *state, const Cand *top, uint32_t outputs, TopState &hs,
                    hipStream_t s) override
    {
        listed_ = nullptr;
        if (outputs > (uint32_t)kMaxOutputs) return Workspace::read_state(d, state, top, outputs, hs, s);
        if (!d->readback) {
            void *p = nullptr;
            check(hipHostMalloc(&p, sizeof(TopState) + kMaxOutputs * sizeof(Cand), hipHostMallocDefault),
                  "hipHostMalloc");
            d->readback = p;
        }
        char *rb = (char *)d->readback;
        check(hipMemcpyAsync(rb, state, sizeof(TopState), hipMemcpyDeviceToHost, s), "state read");
        if (outputs)
            check(hipMemcpyAsync(rb + sizeof(TopState), top, outputs * sizeof(Cand), hipMemcpyDeviceToHost, s),
                  "top read");
        check(hipStreamSynchronize(s), "search sync");
        std::memcpy(&hs, rb, sizeof(hs));
        listed_ = (const Cand *)(rb + sizeof(TopState));
    }
    void results(const Cand *top, uint32_t count, int order, int n, const uint64_t *, int32_t *out_combos,
                 float *out_mi, hipStream_t s) override
    {
        if (listed_) {   // read with the final state
            unpack_results(listed_, count, order, (uint64_t)n, out_combos, out_mi);
            return;
        }
        std::vector<Cand> h(count);
        if (count) check(hipMemcpyAsync(h.data(), top, count * sizeof(Cand), hipMemcpyDeviceToHost, s), "top read");
        check(hipStreamSynchronize(s), "top read sync");
        unpack_results(h.data(), count, order, (uint64_t)n, out_combos, out_mi);
    }

  private:
    const Cand *listed_ = nullptr;
};

// Exhaustive search of `d` in workspace `ws` (fiu_search, fiu_search_static).
void search(fiu_dataset *d, Workspace &ws, int order, int32_t step, int32_t offset, uint32_t outputs,
            const fiu_config *cfg, int32_t *out_combos, float *out_mi, uint32_t *out_count, fiu_search_stats *stats)
{
    const auto t_start = std::chrono::steady_clock::now();
    if (!out_count) fail(FIU_ERR_INVALID, "out_count is NULL");
    *out_count = 0;
    if (order < 2) fail(FIU_ERR_INVALID, "order must be >= 2");
    if (step < 1) fail(FIU_ERR_INVALID, "distribution step must be >= 1");
    if (offset < 0) fail(FIU_ERR_INVALID, "distribution offset must be >= 0");
    if (outputs > 0 && (!out_combos || !out_mi)) fail(FIU_ERR_INVALID, "output pointers are NULL");
    if (order > kMaxOrder) fail(FIU_ERR_UNSUPPORTED, "order " + std::to_string(order) + " is above this build's limit");
    if (outputs > kMaxOutputsBig) fail(FIU_ERR_UNSUPPORTED, "outputs above " + std::to_string(kMaxOutputsBig));
    int algo_req = cfg ? cfg->algorithm : FIU_ALGO_AUTO;
    int variant = cfg && cfg->struct_size >= offsetof(fiu_config, variant) + sizeof(int32_t) ? cfg->variant : 0;
    if (algo_req < FIU_ALGO_AUTO || algo_req > FIU_ALGO_MFMA) fail(FIU_ERR_INVALID, "unknown algorithm");
    if ((algo_req == FIU_ALGO_MARGINAL || algo_req == FIU_ALGO_MFMA) && !d->dev.uniform)
        fail(FIU_ERR_INVALID, "marginal algorithm needs every SNP to cover the same samples");
    if (algo_req == FIU_ALGO_MFMA && order != 2 && order != 3)
        fail(FIU_ERR_UNSUPPORTED, "the matrix-core scan covers orders 2 and 3");
    if (order >= kMinAnyOrder && algo_req == FIU_ALGO_MARGINAL)
        fail(FIU_ERR_UNSUPPORTED, "orders 7 to 12 count every cell (FIU_ALGO_DIRECT)");
    if (algo_req != FIU_ALGO_AUTO && !valid_variant(algo_req, order, variant))
        fail(FIU_ERR_INVALID, "unknown variant " + std::to_string(variant) + " of this algorithm");
    // at least `outputs` candidates per chunk: an overflowing chunk then
    // merges enough entries to set the threshold, and its re-scan pushes only
    // the entries above the N-th best (fewer than N), so re-scans end
    const uint32_t cap = std::max(cfg && cfg->candidate_cap ? cfg->candidate_cap
                                                            : std::min<uint32_t>(1u << 20, ws.max_candidates()),
                                  outputs);
    if (cap > ws.max_candidates())
        fail(FIU_ERR_UNSUPPORTED, "candidate_cap above this entry point's limit of " +
                                      std::to_string(ws.max_candidates()));
    // AUTO: the tuned choice of this shape class when the dataset allows it
    // (order-3 matrix-core entries are timed on whole scans)
    bool from_table = false;
    fiu_tune_entry te{};
    if (algo_req == FIU_ALGO_AUTO && tuned(order, d->snps, (uint64_t)d->cases + d->ctrls, te)) {
        const bool allowed = te.algorithm == FIU_ALGO_DIRECT ||
                             (d->dev.uniform && (te.algorithm != FIU_ALGO_MFMA || order == 2 || step == 1));
        if (allowed) {
            algo_req = te.algorithm;
            if (variant == 0) variant = te.variant;
            from_table = true;
        }
    }
    fiu_search_stats st{};
    st.algorithm = FIU_ALGO_DIRECT;
    const int n = (int)d->snps, M = order - 1;
    if (n < order || outputs == 0) {
        if (stats) *stats = st;
        return;
    }
    // candidates are keyed by their lexicographic rank among the C(n, k)
    // combinations (below the empty marker ~0)
    if (binom_sat((uint64_t)n, (uint64_t)order) >= UINT64_MAX - 1)
        fail(FIU_ERR_UNSUPPORTED, "more than 2^64 - 2 combinations: their ranks do not fit the 64-bit key");
    const uint64_t total_prefixes = binom_sat((uint64_t)n, (uint64_t)M);
    if (total_prefixes == UINT64_MAX) fail(FIU_ERR_UNSUPPORTED, "too many prefixes");
    const uint64_t n_pref = total_prefixes > (uint64_t)offset
                                ? (total_prefixes - (uint64_t)offset + (uint64_t)step - 1) / (uint64_t)step
                                : 0;
    hipStream_t s = stream_of(d, cfg);
    check(hipSetDevice(d->device), "hipSetDevice");
    const uint64_t *binom = ws.binom(d, order, s);   // rows up to k: prefix ranks and combination keys
    int algo = FIU_ALGO_DIRECT;
    const bool small_classes = d->cases < (1u << 24) && d->ctrls < (1u << 24);   // fp32-exact counts
    if (algo_req != FIU_ALGO_DIRECT && d->dev.uniform && order >= 4 && order <= kMaxKmOrder)
        algo = FIU_ALGO_MARGINAL;
    if (algo_req != FIU_ALGO_DIRECT && d->dev.uniform && (order == 2 || order == 3)) {
        algo = FIU_ALGO_MARGINAL;
        if (order == 3 && !ws.pair_table(d, s)) algo = FIU_ALGO_DIRECT;  // table does not fit
        // order 2: the matrix-core scan is the faster exact path (4.4x on
        // 50,000 SNPs x 32,000 samples, also ahead on small inputs)
        if (order == 2 && algo_req != FIU_ALGO_MARGINAL) algo = FIU_ALGO_MFMA;
        // order 3: the matrix-core scan walks every (b, c) of its first SNPs,
        // so by default it takes whole scans (a round-robin share of the
        // prefixes is masked, which wastes the other shares' work)
        if (order == 3 && algo == FIU_ALGO_MARGINAL && small_classes &&
            (algo_req == FIU_ALGO_MFMA || (algo_req == FIU_ALGO_AUTO && step == 1)) && ws.fp4_image(d, s))
            algo = FIU_ALGO_MFMA;
    }
    if (algo != algo_req && from_table) variant = 0;   // the tuned algorithm did not apply
    if (!valid_variant(algo, order, variant)) variant = 0;
    st.algorithm = algo;
    st.variant = variant;
    const int pw = algo == FIU_ALGO_MARGINAL && variant ? variant : kMarginalPW;
    // order-3 matrix-core variants: 3 and 4 the 64 x 32 tile with rings of 3
    // and 4 stages, 5 and 6 the 64 x 64 tile (its counts wait as 16-bit
    // pairs: classes below 2^16 samples, else the 64 x 32 tile).  Default:
    // 6 from 2,048 SNPs and 8,192 samples on (measured: the larger tile is
    // 7 % faster on 10,000 x 16,000, slower on 2,000 x 4,000 and below, where
    // its fewer workgroups and longer epilogue do not pay), else 4
    const uint64_t samples = (uint64_t)d->cases + d->ctrls;
    int ring = algo == FIU_ALGO_MFMA && order == 3 && variant ? variant
                                                              : (n >= 2048 && samples >= 8192 ? 6 : 4);
    if (ring >= 5 && (d->cases >= (1u << 16) || d->ctrls >= (1u << 16) || !ws.private_memory())) ring = 4;
    // matrix-core scans walk super-blocks of 8 x 8 tiles of 64 x 64 pairs;
    // fp4 operands accumulate in fp32, exact while every count is below 2^24;
    // they come from the dataset's fp4 image (int8 when it is not there)
    const int T = d->dev.n_pad / 64;
    const int64_t n_super = mfma2_superblocks(T);
    const bool int8 = (cfg && (cfg->flags & FIU_FLAG_MFMA_INT8)) || (order == 2 && variant == 1);
    const bool mfma_fp4 = algo == FIU_ALGO_MFMA && !int8 && small_classes && ws.fp4_image(d, s);
    // order-2 matrix-core variants: 1 int8 operands, 2 fp4 with the 64 x 64
    // tile, 3 fp4 with the 64 x 128 tile, 4 fp4 with the 128 x 128 tile (the
    // last two: 16-bit counts between the classes, so classes below 2^16
    // samples, else 64 x 64).  Default: 4 from 16,384 SNPs and 8,192 samples
    // on (measured on 50,000 x 32,000: 59.4 ms per scan, 61.2 with 3, 83.6
    // with 2; slower below), else 2
    const bool m2_big = variant == 0 && n >= 16384 && samples >= 8192;
    const int m2_pick = variant == 4 || m2_big ? 2 : variant == 3 ? 1 : 0;
    const int m2_tile = order == 2 && d->cases < (1u << 16) && d->ctrls < (1u << 16) && ws.private_memory()
                            ? m2_pick : 0;

    // chunk plan: even chunks of about `target` combinations (at most
    // max_chunks of them); before the first, a sampling pass over the last
    // prefixes of the distribution (about 2^24 combinations, spread over many
    // short wavefronts) histograms the screening MI and sets a starting
    // threshold, so that the first chunks push few candidates
    const double per_avg = (double)(n - M) / (double)(M + 1);
    const uint64_t target = cfg && cfg->chunk_combos ? cfg->chunk_combos : (1ull << 28);
    const uint64_t max_chunks = ws.max_chunks();
    ChunkPlan plan;
    if (algo == FIU_ALGO_MFMA && order == 2) {
        // chunks of super-blocks (262,144 pairs each) instead of prefixes
        int64_t len = std::max<int64_t>(1, (int64_t)(target / (64 * 4096)));
        // the 64 x 128 tile: launches of at most 64 super-blocks (2,048
        // workgroups, four per slot of the device), whose workgroups on one
        // XCD start close together and so share their row panels in L2 for
        // longer (measured on 50,000 x 32,000: 64.0 ms per scan against 69.9
        // with launches of 1,024 super-blocks; 16 per launch: 70.2, the
        // launch gaps and tails eat the gain)
        if (m2_tile && !(cfg && cfg->chunk_combos)) len = std::min<int64_t>(len, 64);
        len = std::max<int64_t>(len, (int64_t)(((uint64_t)n_super + max_chunks - 1) / max_chunks));
        plan.set_uniform(n_super, len);
    } else if (algo == FIU_ALGO_MFMA) {
        // chunks of first SNPs a, about `target` triples (a, b, c) each; at
        // most 2^36 triples keep a launch's grid within its limit
        uint64_t t = std::min<uint64_t>(target, 1ull << 36);
        const unsigned __int128 all = ChunkPlan::c3(n);
        if (all / t >= max_chunks) t = (uint64_t)(all / max_chunks + 1);
        plan.set_first_snp(n, t);
    } else {
        int64_t len = (int64_t)std::max<double>(1.0, (double)target / std::max(per_avg, 1.0));
        len = std::max<int64_t>(len, (int64_t)((n_pref + max_chunks - 1) / max_chunks));
        plan.set_uniform((int64_t)n_pref, len);
    }
    if (plan.count > max_chunks) fail(FIU_ERR_UNSUPPORTED, "too many launches for this entry point");
    const bool sample = !(cfg && (cfg->flags & FIU_FLAG_NO_THRESHOLD_SAMPLE));
    // at most about 1/16 of the scan's combinations (the pass screens every
    // combination it samples with a popcount kernel), 1/256 ahead of a
    // matrix-core scan, which costs 5-12x less per combination
    const double sample_frac = algo == FIU_ALGO_MFMA ? 1.0 / 256.0 : 1.0 / 16.0;
    // The floor: the scan pushes about outputs x combinations / sample
    // candidates above the starting threshold (twice that with the bins'
    // width), kept near the size of one chunk's candidate buffer; 16,384
    // combinations at least (on 1,000 SNPs x 4,000 samples, top-10: 4,096
    // lets 3x more candidates through the scan and the merge, 65,536 makes
    // the sampling pass cost as much as the scan).
    const double combos_est = (double)n_pref * per_avg;
    const uint64_t sample_floor = (uint64_t)std::max(16384.0, 2.0 * (double)outputs * combos_est / (double)cap);
    const uint64_t sample_cap = std::max<uint64_t>(sample_floor, (uint64_t)(combos_est * sample_frac));
    const int64_t sample_prefixes =
        plan.count == 0 || !sample
            ? 0
            : tail_prefixes(n, M, (uint64_t)step, (uint64_t)offset, total_prefixes, (int64_t)n_pref,
                            std::min<uint64_t>({1ull << 24, target, sample_cap}));
    Cand *cand = nullptr, *top = nullptr;
    TopState *state = nullptr;
    uint32_t *redo = nullptr, *hist = nullptr;
    ws.buffers(d, cap, plan.count, outputs, cand, top, state, redo, hist);
    // more than kMaxOutputs results: the merge runs in global memory
    const bool big_merge = outputs > (uint32_t)kMaxOutputs;
    auto pow2_at_least = [](uint64_t x) {
        uint64_t p = 2;
        while (p < x) p <<= 1;
        return p;
    };
    Cand *mwork = big_merge ? ws.merge_work(d, pow2_at_least((uint64_t)outputs + cap)) : nullptr;
    // state reset and (when sampling) histogram cleared in one launch
    check(launch_search_begin(state, sample_prefixes > 0 ? hist : nullptr, s), "search start launch");

    ScanParams P{};
    P.step = step;
    P.offset = offset;
    P.total_prefixes = total_prefixes;
    P.binom = binom;
    P.cand = cand;
    P.cand_cap = cap;
    P.order = order;
    P.tag = 0;
    P.pair = d->pair_ready ? d->pair.p : nullptr;
    P.cells = order >= kMinBigOrder ? ws.big_cells(d, order) : nullptr;
    mi_constants(d->cases, d->ctrls, d->complete(), P);

    size_t ev = 0;
    bool timed = true;
    auto run_chunk = [&](uint64_t ci, bool time_it) {
        const Chunk c = plan.at(ci);
        if (c.j1 <= c.j0) return;   // an empty first-SNP chunk
        P.j0 = c.j0;
        P.j1 = c.j1;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (time_it && timed) {
            e0 = ws.event(d, ev);
            e1 = ws.event(d, ev + 1);
            timed = e0 && e1;
        }
        if (e0 && e1) check(hipEventRecord(e0, s), "event");
        if (algo == FIU_ALGO_MFMA && order == 2) {
            check(launch_mfma2(d->dev, P, state, c.j0, c.j1, T, mfma_fp4 ? d->fp4.p : nullptr, m2_tile, s),
                  "scan launch");
        } else if (algo == FIU_ALGO_MFMA) {
            check(launch_mfma3(d->dev, P, state, d->fp4.p, (int)c.j0, (int)c.j1, T, ring, s), "scan launch");
        } else if (algo == FIU_ALGO_MARGINAL && order >= 4) {
            check(launch_scan_km(order, d->dev, P, state, s), "scan launch");
        } else if (algo == FIU_ALGO_MARGINAL) {
            check(launch_scan_marginal(order, d->dev, P, state, s, pw), "scan launch");
        } else {
            check(launch_scan_direct(order, d->dev, P, state, s), "scan launch");
        }
        if (e0 && e1) {
            check(hipEventRecord(e1, s), "event");
            ev += 2;
        }
        if (big_merge) {
            TopState h;
            check(hipMemcpyAsync(&h, state, sizeof(h), hipMemcpyDeviceToHost, s), "state read");
            check(hipStreamSynchronize(s), "merge sync");
            const uint32_t count = std::min(h.cand_count, cap);
            check(launch_merge_big(state, top, h.n_top, cand, count, outputs, mwork,
                                   (uint32_t)pow2_at_least((uint64_t)h.n_top + count), (uint32_t)ci, redo, s),
                  "merge launch");
        } else {
            check(launch_merge(state, top, cand, outputs, cap, (uint32_t)ci, redo, s), "merge launch");
        }
        ++st.launches;
    };
    if (sample_prefixes > 0) {
        ScanParams H = P;
        H.j0 = (int64_t)n_pref - sample_prefixes;
        H.j1 = (int64_t)n_pref;
        H.hist = hist;
        if (algo == FIU_ALGO_MARGINAL && order >= 4)
            check(launch_scan_km(order, d->dev, H, state, s), "sample launch");
        else if (algo == FIU_ALGO_MARGINAL || algo == FIU_ALGO_MFMA)
            check(launch_scan_marginal(order, d->dev, H, state, s), "sample launch");
        else
            check(launch_scan_direct(order, d->dev, H, state, s), "sample launch");
        check(launch_tau_from_hist(hist, outputs, P.delta, state, s), "threshold launch");
    }
    for (uint64_t ci = 0; ci < plan.count; ++ci) run_chunk(ci, true);
    // chunks whose candidates overflowed the buffer are scanned again with
    // the (higher) final threshold until none overflows
    TopState hs;
    for (;;) {
        ws.read_state(d, state, top, outputs, hs, s);
        if (hs.redo_count == 0) break;
        const uint32_t n_redo = hs.redo_count;
        hs.redo_count = 0;
        check(hipMemcpyAsync(state, &hs, sizeof(hs), hipMemcpyHostToDevice, s), "state reset");
        // the list is read in blocks of 1024, each block's re-scans queued
        // before the next block is read.  Re-scans append to the list from
        // entry 0 again, one entry at most per re-scan, so the re-scans of
        // blocks 0 .. b write below entry 1024 (b + 1): the unread blocks
        // are intact when their (stream-ordered) read runs.
        uint32_t ids[1024];
        for (uint32_t b0 = 0; b0 < n_redo; b0 += 1024) {
            const uint32_t nb = std::min<uint32_t>(1024, n_redo - b0);
            check(hipMemcpyAsync(ids, redo + b0, nb * 4, hipMemcpyDeviceToHost, s), "redo read");
            check(hipStreamSynchronize(s), "redo read sync");
            for (uint32_t i = 0; i < nb; ++i) {
                run_chunk(ids[i], false);
                ++st.redo_chunks;
            }
        }
    }
    double scan_ms = 0.0;
    for (size_t i = 0; i + 1 < ev; i += 2) {
        float ms = 0.0f;
        check(hipEventElapsedTime(&ms, ws.event(d, i), ws.event(d, i + 1)), "event time");
        scan_ms += ms;
    }
    const uint32_t n_top = std::min<uint32_t>(hs.n_top, outputs);
    ws.results(top, n_top, order, n, binom, out_combos, out_mi, s);
    *out_count = n_top;
    // combinations evaluated: closed form for the whole space, else a walk
    // over the runs of prefixes (O(C(n, order-2)))
    st.combinations = (step == 1 && offset == 0) ? binom_sat((uint64_t)n, (uint64_t)order)
                                                 : count_combinations((uint64_t)n, M, (uint64_t)step, (uint64_t)offset);
    st.scan_ms = scan_ms;
    st.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    if (stats) *stats = st;
}

// ---------------------------------------------------------------------------
// fiu_search_static: the same search in one fixed block of device memory.
// ---------------------------------------------------------------------------
// Layout of the block (each region 256-byte aligned).  The limits it sets:
// 2^18 candidates per chunk, 2^14 launches per search, 2^19 binomial entries
// ((order + 1) x (snps + 1)).
constexpr size_t kStaticCand = 1u << 18;
constexpr size_t kStaticChunks = 1u << 14;
constexpr size_t kStaticBinom = 1u << 19;
constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t kOffState = 0;
constexpr size_t kOffTop = kOffState + align256(sizeof(TopState));
constexpr size_t kOffCand = kOffTop + align256(kMaxOutputs * sizeof(Cand));
constexpr size_t kOffRedo = kOffCand + align256(kStaticCand * sizeof(Cand));
constexpr size_t kOffHist = kOffRedo + align256(kStaticChunks * sizeof(uint32_t));
constexpr size_t kOffCombos = kOffHist + align256(kHistBins * sizeof(uint32_t));
constexpr size_t kOffMi = kOffCombos + align256((size_t)kMaxOutputs * kMaxOrder * sizeof(int32_t));
constexpr size_t kOffBinom = kOffMi + align256(kMaxOutputs * sizeof(float));
constexpr size_t kStaticBytes = kOffBinom + align256(kStaticBinom * sizeof(uint64_t));
static_assert(kStaticBytes < (9u << 20), "the fixed block stays below 9 MiB");

}  // namespace

// The fixed block (one copy per GPU, held until the process ends).  In host
// code this name is not a device address: every host use goes through the
// address hipGetSymbolAddress returns for the current device (static_block).
static __device__ __align__(256) unsigned char scratch[kStaticBytes];

namespace {

constexpr int kStaticDevices = 64;
std::mutex g_static_mu;
void *g_static_ptr[kStaticDevices];   // per device, filled at the first call on it

// The fixed block of `device` (current): its device address, looked up once
// per device with that device current.
char *static_block(int device)
{
    if (device < 0 || device >= kStaticDevices) fail(FIU_ERR_UNSUPPORTED, "device index above 63");
    std::lock_guard<std::mutex> lk(g_static_mu);
    if (g_static_ptr[device]) return (char *)g_static_ptr[device];
    void *ptr = nullptr;
    check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!ptr) fail(FIU_ERR_DEVICE, "no device address for the fixed search block");
    g_static_ptr[device] = ptr;
    return (char *)ptr;
}

class StaticWorkspace final : public Workspace {
  public:
    explicit StaticWorkspace(char *base) : base_(base) {}
    const uint64_t *binom(fiu_dataset *d, int m, hipStream_t s) override
    {
        const uint64_t entries = (uint64_t)(m + 1) * ((uint64_t)d->snps + 1);
        if (entries > kStaticBinom)
            fail(FIU_ERR_UNSUPPORTED, "fiu_search_static: (order + 1) x (snps + 1) = " + std::to_string(entries) +
                                          " binomial entries, above its limit of " + std::to_string(kStaticBinom));
        uint64_t *b = (uint64_t *)(base_ + kOffBinom);
        check(launch_binom_table(b, d->snps, m, s), "binomial table launch");
        return b;
    }
    // derived tables an earlier fiu_search built; none is built here
    bool pair_table(fiu_dataset *d, hipStream_t) override
    {
        std::lock_guard<std::mutex> lk(d->prep_mu);
        return d->pair_ready;
    }
    bool fp4_image(fiu_dataset *d, hipStream_t) override
    {
        std::lock_guard<std::mutex> lk(d->prep_mu);
        return d->fp4_ready;
    }
    uint32_t max_candidates() const override { return (uint32_t)kStaticCand; }
    uint64_t max_chunks() const override { return kStaticChunks; }
    // the tiles without private memory: no launch of this search needs more
    // than the fixed block and the caller's arrays
    bool private_memory() const override { return false; }
    void buffers(fiu_dataset *, uint32_t cap, uint64_t chunks, uint32_t outputs, Cand *&cand, Cand *&top,
                 TopState *&state, uint32_t *&redo, uint32_t *&hist) override
    {
        if (cap > kStaticCand || chunks > kStaticChunks) fail(FIU_ERR_UNSUPPORTED, "fiu_search_static: limits exceeded");
        if (outputs > (uint32_t)kMaxOutputs)
            fail(FIU_ERR_UNSUPPORTED, "fiu_search_static: outputs above " + std::to_string(kMaxOutputs));
        cand = (Cand *)(base_ + kOffCand);
        top = (Cand *)(base_ + kOffTop);
        state = (TopState *)(base_ + kOffState);
        redo = (uint32_t *)(base_ + kOffRedo);
        hist = (uint32_t *)(base_ + kOffHist);
    }
    uint32_t *big_cells(fiu_dataset *, int order) override
    {
        fail(FIU_ERR_UNSUPPORTED, "fiu_search_static: orders above 9 need " + std::to_string(big_cells_bytes(order)) +
                                      " bytes of counters, beyond its fixed block");
    }
    Cand *merge_work(fiu_dataset *, size_t) override
    {
        fail(FIU_ERR_UNSUPPORTED, "fiu_search_static: outputs above " + std::to_string(kMaxOutputs));
    }
    hipEvent_t event(fiu_dataset *, size_t) override { return nullptr; }
    void results(const Cand *top, uint32_t count, int order, int n, const uint64_t *binom, int32_t *out_combos,
                 float *out_mi, hipStream_t s) override
    {
        int32_t *combos = (int32_t *)(base_ + kOffCombos);
        float *mi = (float *)(base_ + kOffMi);
        check(launch_unpack_top(top, count, order, n, binom, combos, mi, s), "unpack launch");
        if (count) {
            check(hipMemcpyAsync(out_combos, combos, (size_t)count * order * sizeof(int32_t), hipMemcpyDeviceToHost, s),
                  "results read");
            check(hipMemcpyAsync(out_mi, mi, (size_t)count * sizeof(float), hipMemcpyDeviceToHost, s), "results read");
        }
        check(hipStreamSynchronize(s), "results sync");
    }

  private:
    char *base_;
};

}  // namespace

namespace fiu {

void unpack_results(const Cand *top, uint32_t count, int order, uint64_t n, int32_t *out_combos, float *out_mi)
{
    for (uint32_t i = 0; i < count; ++i) {
        unrank_host(top[i].combo, n, order, out_combos + (size_t)i * order);
        out_mi[i] = top[i].mi;
    }
}

}  // namespace fiu

extern "C" {

const char *fiu_last_error(void) { return error_slot().c_str(); }

const char *fiu_version(void) { return "fiuncho_amd 0.2 (gfx950)"; }

fiu_status fiu_set_device(int device)
{
    return guard([&] {
        int count = 0;
        check(hipGetDeviceCount(&count), "hipGetDeviceCount");
        if (device < 0 || device >= count) fail(FIU_ERR_INVALID, "no such device");
        current_device() = device;
    });
}

fiu_status fiu_set_tuning(const fiu_tune_entry *entries, size_t n)
{
    return guard([&] {
        if (n && !entries) fail(FIU_ERR_INVALID, "entries is NULL");
        for (size_t i = 0; i < n; ++i) {
            const fiu_tune_entry &e = entries[i];
            if (e.order < 2 || e.order > kMaxOrder) fail(FIU_ERR_INVALID, "tuning entry with an unsupported order");
            if (e.algorithm < FIU_ALGO_DIRECT || e.algorithm > FIU_ALGO_MFMA)
                fail(FIU_ERR_INVALID, "tuning entry with an unknown algorithm");
            if (e.algorithm == FIU_ALGO_MFMA && e.order != 2 && e.order != 3)
                fail(FIU_ERR_INVALID, "matrix-core tuning entry above order 3");
            if (e.algorithm == FIU_ALGO_MARGINAL && e.order >= kMinAnyOrder)
                fail(FIU_ERR_INVALID, "marginal tuning entry above order 6");
            if (!valid_variant(e.algorithm, e.order, e.variant)) fail(FIU_ERR_INVALID, "tuning entry with an unknown variant");
            if (e.snps_lo >= e.snps_hi || e.samples_lo >= e.samples_hi)
                fail(FIU_ERR_INVALID, "tuning entry with an empty range");
            if (e.batch_unit > 6) fail(FIU_ERR_INVALID, "tuning entry with a batch unit above 6");
            if (e.algorithm == FIU_ALGO_MFMA && e.batch_unit > 2)
                fail(FIU_ERR_INVALID, "matrix-core tuning entry with a batch tile above 2");
        }
        std::lock_guard<std::mutex> lk(g_tuning_mu);
        g_tuning.assign(entries, entries + n);
    });
}

fiu_status fiu_get_tuning(fiu_tune_entry *entries, size_t cap, size_t *n)
{
    return guard([&] {
        if (!n) fail(FIU_ERR_INVALID, "n is NULL");
        if (cap && !entries) fail(FIU_ERR_INVALID, "entries is NULL");
        std::lock_guard<std::mutex> lk(g_tuning_mu);
        *n = g_tuning.size();
        std::copy(g_tuning.begin(), g_tuning.begin() + std::min(cap, g_tuning.size()), entries);
    });
}

void fiu_config_init(fiu_config *cfg)
{
    if (!cfg) return;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->struct_size = sizeof(fiu_config);
    cfg->algorithm = FIU_ALGO_AUTO;
}

fiu_status fiu_dataset_create(const uint64_t *words, size_t snps, size_t cases_words, size_t ctrls_words,
                              uint32_t cases, uint32_t ctrls, fiu_dataset **out)
{
    return guard([&] {
        if (!out) fail(FIU_ERR_INVALID, "out is NULL");
        *out = create_dataset(words, snps, cases_words, ctrls_words, cases, ctrls);
    });
}

fiu_status fiu_dataset_read(const char *const *paths, size_t npaths, fiu_dataset **out)
{
    return guard([&] {
        if (!out || (!paths && npaths)) fail(FIU_ERR_INVALID, "NULL argument");
        std::vector<std::string> p;
        for (size_t i = 0; i < npaths; ++i) p.emplace_back(paths[i] ? paths[i] : "");
        Genotypes g;
        parse_files(p, g);
        *out = dataset_from_genotypes(g.geno.data(), g.snps, g.samples, g.is_ctrl.data(), true);
    });
}

fiu_status fiu_dataset_from_genotypes(const int8_t *geno, size_t snps, size_t samples, const uint8_t *is_ctrl,
                                      uint32_t flags, fiu_dataset **out)
{
    return guard([&] {
        if (!out) fail(FIU_ERR_INVALID, "out is NULL");
        if (flags & ~FIU_GENO_READER_FLUSH) fail(FIU_ERR_INVALID, "unknown flags");
        *out = dataset_from_genotypes(geno, snps, samples, is_ctrl, (flags & FIU_GENO_READER_FLUSH) != 0);
    });
}

fiu_status fiu_read_files(const char *const *paths, size_t npaths, uint64_t *words, size_t capacity, size_t *snps,
                          uint32_t *cases, uint32_t *ctrls, size_t *cases_words, size_t *ctrls_words)
{
    return guard([&] {
        if (!paths && npaths) fail(FIU_ERR_INVALID, "paths is NULL");
        std::vector<std::string> p;
        for (size_t i = 0; i < npaths; ++i) p.emplace_back(paths[i] ? paths[i] : "");
        PackedDataset pk;
        read_files(p, pk);
        if (snps) *snps = pk.snps;
        if (cases) *cases = (uint32_t)pk.cases;
        if (ctrls) *ctrls = (uint32_t)pk.ctrls;
        if (cases_words) *cases_words = pk.cases_words;
        if (ctrls_words) *ctrls_words = pk.ctrls_words;
        if (words) {
            if (capacity < pk.words.size()) fail(FIU_ERR_INVALID, "words buffer too small");
            std::memcpy(words, pk.words.data(), pk.words.size() * 8);
        }
    });
}

fiu_status fiu_dataset_info(const fiu_dataset *ds, size_t *snps, uint32_t *cases, uint32_t *ctrls,
                            size_t *cases_words, size_t *ctrls_words, size_t *device_bytes)
{
    return guard([&] {
        if (!ds) fail(FIU_ERR_INVALID, "dataset is NULL");
        if (snps) *snps = ds->snps;
        if (cases) *cases = ds->cases;
        if (ctrls) *ctrls = ds->ctrls;
        if (cases_words) *cases_words = ds->cw;
        if (ctrls_words) *ctrls_words = ds->tw;
        if (device_bytes) *device_bytes = ds->bits.n * sizeof(uint4) + ds->singles.n * 4;
    });
}

fiu_status fiu_dataset_words(const fiu_dataset *ds, uint64_t *out)
{
    return guard([&] {
        if (!ds || !out) fail(FIU_ERR_INVALID, "NULL argument");
        std::memcpy(out, ds->host_words.data(), ds->host_words.size() * 8);
    });
}

fiu_status fiu_dataset_destroy(fiu_dataset *ds)
{
    return guard([&] {
        if (ds) {
            (void)hipSetDevice(ds->device);
            if (ds->stream) (void)hipStreamSynchronize(ds->stream);
        }
        delete ds;
    });
}

fiu_status fiu_search(const fiu_dataset *ds, uint16_t order, int32_t dist_step, int32_t dist_offset,
                      uint32_t outputs, const fiu_config *cfg, int32_t *out_combos, float *out_mi,
                      uint32_t *out_count, fiu_search_stats *stats)
{
    return guard([&] {
        if (!ds) fail(FIU_ERR_INVALID, "dataset is NULL");
        if (cfg && cfg->struct_size != sizeof(fiu_config)) fail(FIU_ERR_INVALID, "fiu_config.struct_size mismatch");
        DatasetWorkspace ws;
        search(const_cast<fiu_dataset *>(ds), ws, order, dist_step, dist_offset, outputs, cfg, out_combos, out_mi,
               out_count, stats);
    });
}

fiu_status fiu_search_static(const fiu_dataset *ds, uint16_t order, int32_t dist_step, int32_t dist_offset,
                             uint32_t outputs, const fiu_config *cfg, int32_t *out_combos, float *out_mi,
                             uint32_t *out_count, fiu_search_stats *stats)
{
    return guard([&] {
        if (!ds) fail(FIU_ERR_INVALID, "dataset is NULL");
        if (cfg && cfg->struct_size != sizeof(fiu_config)) fail(FIU_ERR_INVALID, "fiu_config.struct_size mismatch");
        check(hipSetDevice(ds->device), "hipSetDevice");
        StaticWorkspace ws(static_block(ds->device));
        search(const_cast<fiu_dataset *>(ds), ws, order, dist_step, dist_offset, outputs, cfg, out_combos, out_mi,
               out_count, stats);
    });
}

size_t fiu_search_static_bytes(void) { return kStaticBytes; }

size_t fiu_ct_size(uint16_t order)
{
    size_t s = 1;
    for (int i = 0; i < order; ++i) s *= 3;
    return (s + 15) / 16 * 16;
}

static void check_combos(const fiu_dataset *ds, const int32_t *combos, int order, size_t n)
{
    if (!combos && n) fail(FIU_ERR_INVALID, "combos is NULL");
    for (size_t i = 0; i < n; ++i)
        for (int q = 0; q < order; ++q) {
            const int32_t c = combos[i * order + q];
            if (c < 0 || (size_t)c >= ds->snps) fail(FIU_ERR_INVALID, "SNP index out of range");
            if (q && c <= combos[i * order + q - 1]) fail(FIU_ERR_INVALID, "combination not strictly ascending");
        }
}

fiu_status fiu_combine_batch(const fiu_dataset *ds, const int32_t *combos, uint16_t order, size_t n, uint64_t *out)
{
    return guard([&] {
        if (!ds || (!out && n)) fail(FIU_ERR_INVALID, "NULL argument");
        if (order < 1 || order > 12) fail(FIU_ERR_UNSUPPORTED, "order outside 1..12");
        check_combos(ds, combos, order, n);
        if (!n) return;
        check(hipSetDevice(ds->device), "hipSetDevice");
        uint64_t R = 1;
        for (int i = 0; i < order; ++i) R *= 3;
        const size_t per = R * (ds->cw + ds->tw);
        DBuf<int32_t> dc;
        DBuf<uint64_t> dout;
        dc.alloc(n * order);
        dout.alloc(n * per);
        check(hipMemcpyAsync(dc.p, combos, n * order * 4, hipMemcpyHostToDevice, ds->stream), "copy");
        check(launch_combine_batch(ds->dev, dc.p, order, n, ds->cw, ds->tw, dout.p, ds->stream), "combine");
        check(hipMemcpyAsync(out, dout.p, n * per * 8, hipMemcpyDeviceToHost, ds->stream), "copy back");
        check(hipStreamSynchronize(ds->stream), "sync");
    });
}

fiu_status fiu_contingency_batch(const fiu_dataset *ds, const int32_t *combos, uint16_t order, size_t n,
                                 uint32_t *out_counts, float *out_mi)
{
    return guard([&] {
        if (!ds || (!out_counts && n)) fail(FIU_ERR_INVALID, "NULL argument");
        if (order < 1 || order > kMaxOrder) fail(FIU_ERR_UNSUPPORTED, "order outside 1..12");
        check_combos(ds, combos, order, n);
        if (!n) return;
        check(hipSetDevice(ds->device), "hipSetDevice");
        const size_t cts = fiu_ct_size(order);
        DBuf<int32_t> dc;
        DBuf<uint32_t> dout;
        DBuf<float> dmi;
        dc.alloc(n * order);
        dout.alloc(n * 2 * cts);
        check(hipMemcpyAsync(dc.p, combos, n * order * 4, hipMemcpyHostToDevice, ds->stream), "copy");
        check(launch_contingency_batch(ds->dev, dc.p, order, n, cts, dout.p, ds->stream), "contingency");
        if (out_mi) {
            ScanParams P{};
            mi_constants(ds->cases, ds->ctrls, false, P);
            dmi.alloc(n);
            check(launch_mi_batch(dout.p, n, cts, P.n64, P.hy64, dmi.p, ds->stream), "mi");
            check(hipMemcpyAsync(out_mi, dmi.p, n * 4, hipMemcpyDeviceToHost, ds->stream), "copy back");
        }
        check(hipMemcpyAsync(out_counts, dout.p, n * 2 * cts * 4, hipMemcpyDeviceToHost, ds->stream), "copy back");
        check(hipStreamSynchronize(ds->stream), "sync");
    });
}

fiu_status fiu_mi_batch(const uint32_t *tables, size_t n, uint16_t order, uint32_t cases, uint32_t ctrls,
                        float *out_mi)
{
    return guard([&] {
        if ((!tables || !out_mi) && n) fail(FIU_ERR_INVALID, "NULL argument");
        if (order < 1 || order > 12) fail(FIU_ERR_UNSUPPORTED, "order outside 1..12");
        if (!n) return;
        check(hipSetDevice(current_device()), "hipSetDevice");
        const size_t cts = fiu_ct_size(order);
        hipStream_t s;
        check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream");
        std::unique_ptr<void, void (*)(void *)> sg((void *)s, [](void *p) { (void)hipStreamDestroy((hipStream_t)p); });
        DBuf<uint32_t> dt;
        DBuf<float> dm;
        dt.alloc(n * 2 * cts);
        dm.alloc(n);
        check(hipMemcpyAsync(dt.p, tables, n * 2 * cts * 4, hipMemcpyHostToDevice, s), "copy");
        ScanParams P{};
        mi_constants(cases, ctrls, false, P);
        check(launch_mi_batch(dt.p, n, cts, P.n64, P.hy64, dm.p, s), "mi");
        check(hipMemcpyAsync(out_mi, dm.p, n * 4, hipMemcpyDeviceToHost, s), "copy back");
        check(hipStreamSynchronize(s), "sync");
    });
}

static SynthSpec to_spec(const fiu_synth_spec *sp)
{
    if (!sp) fail(FIU_ERR_INVALID, "spec is NULL");
    if (sp->struct_size != sizeof(fiu_synth_spec)) fail(FIU_ERR_INVALID, "fiu_synth_spec.struct_size mismatch");
    if (sp->planted_order > 8) fail(FIU_ERR_INVALID, "planted_order above 8");
    SynthSpec s;
    s.snps = sp->snps;
    s.cases = sp->cases;
    s.ctrls = sp->ctrls;
    s.maf_lo = sp->maf_lo;
    s.maf_hi = sp->maf_hi;
    s.low_frac = sp->low_frac;
    s.low_lo = sp->low_lo;
    s.low_hi = sp->low_hi;
    s.seed = sp->seed;
    s.planted_order = (int)sp->planted_order;
    for (int i = 0; i < 8; ++i) s.planted[i] = sp->planted[i];
    for (uint32_t i = 0; i < sp->planted_order; ++i) {
        if (sp->planted[i] < 0 || (uint64_t)sp->planted[i] >= sp->snps) fail(FIU_ERR_INVALID, "planted SNP out of range");
        if (i && sp->planted[i] <= sp->planted[i - 1]) fail(FIU_ERR_INVALID, "planted SNPs not ascending");
    }
    s.pen_seed = sp->penetrance_seed;
    s.pen_lo = sp->pen_lo;
    s.pen_hi = sp->pen_hi;
    return s;
}

static int synth_threads(const fiu_synth_spec *sp)
{
    return sp->threads > 0 ? sp->threads : (int)std::max(1u, std::thread::hardware_concurrency());
}

void fiu_synth_spec_init(fiu_synth_spec *spec)
{
    if (!spec) return;
    std::memset(spec, 0, sizeof(*spec));
    spec->struct_size = sizeof(fiu_synth_spec);
    spec->maf_lo = 0.05;
    spec->maf_hi = 0.5;
    spec->low_lo = 0.01;
    spec->low_hi = 0.05;
    spec->seed = 1;
    spec->penetrance_seed = 2;
    spec->pen_lo = 0.3;
    spec->pen_hi = 0.7;
}

fiu_status fiu_synth_generate(const fiu_synth_spec *spec, uint64_t *words, int8_t *geno, size_t *cases_words,
                              size_t *ctrls_words)
{
    return guard([&] {
        const SynthSpec s = to_spec(spec);
        if (cases_words) *cases_words = synth_words(s.cases);
        if (ctrls_words) *ctrls_words = synth_words(s.ctrls);
        if (words || geno) synth_generate(s, words, geno, synth_threads(spec));
    });
}

fiu_status fiu_dataset_synth(const fiu_synth_spec *spec, fiu_dataset **out)
{
    return guard([&] {
        if (!out) fail(FIU_ERR_INVALID, "out is NULL");
        const SynthSpec s = to_spec(spec);
        const size_t cw = synth_words(s.cases), tw = synth_words(s.ctrls);
        std::vector<uint64_t> words(s.snps * 3 * (cw + tw));
        synth_generate(s, words.data(), nullptr, synth_threads(spec));
        *out = create_dataset(words.data(), s.snps, cw, tw, s.cases, s.ctrls);
    });
}

}  // extern "C"

// Batched scans and the scan server.
//
// Many order-k searches advance together through batched steps.  A step is
// one fixed HIP graph: a persistent scan launch that works through a slice of
// every active scan (whatever their datasets, sizes and orders), a second
// persistent launch for order-4 scans, and one top-N merge per scan.  Only
// the step descriptor (which prefix ranges of which scans) changes from step
// to step; it is copied to the device before the graph is launched.  The host
// keeps two steps in flight, so the device never waits for planning.
//
// Per scan, the first slice is small (the exact top-N of its first
// combinations sets the threshold) and every later slice is at most as large
// as everything given to that scan before, so the candidate buffers stay
// small; a slice whose candidates overflowed is scanned again.  Results do not
// depend on the slicing: counts are exact, the MI is computed in FP64 from the
// counts, and ties are broken by the combination.
//
// The scan server (fiu_server_*) runs the same engine on its own thread and
// adds admission: requests carrying host tables are uploaded into one device
// memory budget (an Arena) when their footprint fits, by an ingest thread on
// its own stream, while the scheduler thread keeps stepping the scans already
// on the device.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "fiu_internal.h"

using namespace fiu;

namespace fiu {

// ---------------------------------------------------------------------------
// Arena
// ---------------------------------------------------------------------------
Arena::Arena(size_t bytes, int device, bool pinned_host) : pinned_(pinned_host)
{
    bytes = (std::max<size_t>(bytes, kAlign) + kAlign - 1) / kAlign * kAlign;
    check(hipSetDevice(device), "hipSetDevice");
    const hipError_t e = pinned_ ? hipHostMalloc((void **)&base_, bytes, hipHostMallocDefault)
                                 : hipMalloc((void **)&base_, bytes);
    if (e != hipSuccess) {
        base_ = nullptr;
        (void)hipGetLastError();
        fail(FIU_ERR_NOMEM, std::string("scan server: cannot allocate ") + (pinned_ ? "pinned host" : "device") +
                                " memory of " + std::to_string(bytes) + " bytes");
    }
    size_ = bytes;
    free_[0] = bytes;
}

Arena::~Arena()
{
    if (base_) {
        if (pinned_) (void)hipHostFree(base_);
        else (void)hipFree(base_);
    }
}

void *Arena::alloc(size_t bytes)
{
    bytes = (std::max<size_t>(bytes, 1) + kAlign - 1) / kAlign * kAlign;
    std::lock_guard<std::mutex> lk(mu_);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second < bytes) continue;
        const size_t off = it->first, len = it->second;
        free_.erase(it);
        if (len > bytes) free_[off + bytes] = len - bytes;
        taken_[off] = bytes;
        used_ += bytes;
        return base_ + off;
    }
    return nullptr;
}

void Arena::release(void *p)
{
    std::lock_guard<std::mutex> lk(mu_);
    size_t off = (size_t)((char *)p - base_);
    auto t = taken_.find(off);
    if (t == taken_.end()) return;
    size_t len = t->second;
    taken_.erase(t);
    used_ -= len;
    auto nx = free_.lower_bound(off);
    if (nx != free_.end() && nx->first == off + len) {
        len += nx->second;
        nx = free_.erase(nx);
    }
    if (nx != free_.begin()) {
        auto pv = std::prev(nx);
        if (pv->first + pv->second == off) {
            pv->second += len;
            return;
        }
    }
    free_[off] = len;
}

size_t Arena::used() const
{
    std::lock_guard<std::mutex> lk(mu_);
    return used_;
}

size_t Arena::largest_free() const
{
    std::lock_guard<std::mutex> lk(mu_);
    size_t m = 0;
    for (const auto &f : free_) m = std::max(m, f.second);
    return m;
}

}  // namespace fiu

namespace {

using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point t0) { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }

// Host state of one scan.
struct Job {
    uint64_t id = 0;
    fiu_dataset *ds = nullptr;
    std::unique_ptr<fiu_dataset> owned;   // server-uploaded dataset
    std::vector<uint64_t> words;          // server: host tables until upload
    uint64_t snps = 0, cw = 0, tw = 0;
    uint32_t cases = 0, ctrls = 0;
    bool staged = false;                  // device work of the upload still to enqueue
    StagedUpload up;
    const uint64_t *binom = nullptr;      // binomial table of the dataset for this order
    int order = 0;
    int32_t step = 1, offset = 0;
    uint32_t outputs = 0;
    // plan
    int kind = kKindDirect2, pw = 1, chunks = 1, tiles = 1;
    bool wide = false;   // matrix-core kinds: the 64 x 64 / 64 x 128 tiles
    // units of work: prefixes (popcount kinds), first SNPs (order 3 on the
    // matrix cores) or super-blocks (order 2 on the matrix cores)
    int64_t n_pref = 0, next = 0, given = 0, seed = 1;
    double instr_per_prefix = 1.0;    // popcount kinds: cost of one prefix
    double instr_per_combo = 0.0;     // matrix-core kinds: cost of one combination
    std::deque<std::pair<int64_t, int64_t>> redo;
    int inflight = 0;
    int slot = -1;
    // result
    fiu_status status = FIU_OK;
    std::string error;
    std::vector<int32_t> combos;
    std::vector<float> mi;
    uint32_t count = 0;
    fiu_request_stats rs{};
    bool done = false;
    fiu_scan *scan = nullptr;             // fiu_search_batch: where results go
};

struct EngineConfig {
    int max_active = 128;
    double step_ms = 6.0;
    uint32_t cand_cap = 1u << 16;
};

template <class T> T *pinned(size_t n)
{
    void *p = nullptr;
    check(hipHostMalloc(&p, std::max<size_t>(n, 1) * sizeof(T), hipHostMallocDefault), "hipHostMalloc");
    return (T *)p;
}

// Scheduler core: owns the per-slot device buffers, the step graph and the
// step pipeline.  All members except ready() / wake() are used by one thread.
class Engine {
  public:
    Engine(int device, const EngineConfig &cfg, Clock::time_point epoch);
    ~Engine();

    // Queue a prepared job (dataset resident, binom / pair tables built).
    void ready(Job *j);
    // One scheduling round: activate queued jobs into free slots, launch a
    // step if there is work, complete the oldest step when two are in flight
    // (or when nothing new could be launched).  Returns false when idle.
    bool pump();
    bool idle();
    void set_on_done(std::function<void(Job *)> f) { on_done_ = std::move(f); }
    // The stagings of staged uploads (pinned host block, device block) are
    // released here once the step after their copies has completed; `freed`
    // runs after each such release.
    void set_staging(Arena *pinned, std::function<void()> freed)
    {
        pinned_ = pinned;
        on_staging_freed_ = std::move(freed);
    }
    fiu_batch_stats stats() const { return stats_; }
    hipStream_t prep_stream() const { return prep_; }
    // Plan constants of a job (kind, prefixes, unit shape); throws on
    // invalid or unsupported requests.
    static void plan_job(Job *j);

  private:
    struct StepRec {
        std::vector<std::tuple<Job *, int64_t, int64_t>> ranges;
        std::vector<Job *> jobs;
        double est_instr = 0.0;
        bool valid = false;
    };

    void activate(Job *j, int slot);
    bool launch_step();
    void complete_step(int parity);
    void finish(Job *j, uint32_t n_top);

    int device_;
    EngineConfig cfg_;
    Clock::time_point epoch_;
    hipStream_t stream_ = nullptr, rb_ = nullptr, prep_ = nullptr;
    DBuf<BatchJob> jobs_d_;
    DBuf<TopState> states_d_;
    DBuf<Cand> cand_d_, top_d_;
    DBuf<char> desc_d_;
    BatchJob *jobs_h_ = nullptr;
    TopState *init_h_ = nullptr;
    char *desc_h_[2] = {nullptr, nullptr};
    TopState *states_h_[2] = {nullptr, nullptr};
    Cand *top_h_ = nullptr;   // pinned readback of one finished scan's top list
    hipEvent_t ev_start_[2] = {}, ev_end_[2] = {}, ev_done_[2] = {};
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
    size_t max_segs_ = 0, max_msegs_ = 0, off_merge_ = 0, off_segs0_ = 0, off_segs1_ = 0, off_segs2_ = 0,
           off_segs3_ = 0, desc_bytes_ = 0;
    std::vector<Job *> slots_;
    std::vector<Job *> active_;
    std::mutex ready_mu_;
    std::deque<Job *> ready_;
    StepRec rec_[2];
    uint64_t step_no_ = 0;   // steps launched
    uint64_t done_no_ = 0;   // steps completed
    double rate_ = 8.0e8;    // VALU wave-instructions per ms (calibrated)
    fiu_batch_stats stats_{};
    std::function<void(Job *)> on_done_;
    Arena *pinned_ = nullptr;
    std::function<void()> on_staging_freed_;
    struct Deferred {
        uint64_t step;   // released once this step has completed
        void *block;
        Arena *arena;
    };
    std::vector<Deferred> deferred_;
};

Engine::Engine(int device, const EngineConfig &cfg, Clock::time_point epoch) : device_(device), cfg_(cfg), epoch_(epoch)
{
    check(hipSetDevice(device_), "hipSetDevice");
    check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    check(hipStreamCreateWithFlags(&rb_, hipStreamNonBlocking), "hipStreamCreate");
    check(hipStreamCreateWithFlags(&prep_, hipStreamNonBlocking), "hipStreamCreate");
    const int A = cfg_.max_active;
    max_segs_ = (size_t)A * 8 + 16;
    off_merge_ = 256;
    off_segs0_ = (off_merge_ + (size_t)A * 4 + 255) / 256 * 256;
    off_segs1_ = off_segs0_ + max_segs_ * sizeof(BatchSeg);
    // matrix-core segments: order-3 ranges are cut at every 64th first SNP
    max_msegs_ = (size_t)A * 64 + 1024;
    off_segs2_ = off_segs1_ + max_segs_ * sizeof(BatchSeg);
    off_segs3_ = off_segs2_ + max_msegs_ * sizeof(BatchSeg);
    desc_bytes_ = off_segs3_ + max_segs_ * sizeof(BatchSeg);
    jobs_d_.alloc(A);
    states_d_.alloc(A);
    cand_d_.alloc((size_t)A * cfg_.cand_cap);
    top_d_.alloc((size_t)A * kMaxOutputs);
    desc_d_.alloc(desc_bytes_);
    jobs_h_ = pinned<BatchJob>(A);
    top_h_ = pinned<Cand>(kMaxOutputs);
    init_h_ = pinned<TopState>(1);
    std::memset(init_h_, 0, sizeof(TopState));
    init_h_->tau = -INFINITY;
    init_h_->tau_combo = ~0ull;
    for (int p = 0; p < 2; ++p) {
        desc_h_[p] = pinned<char>(desc_bytes_);
        std::memset(desc_h_[p], 0, desc_bytes_);
        states_h_[p] = pinned<TopState>(A);
        check(hipEventCreate(&ev_start_[p]), "hipEventCreate");
        check(hipEventCreate(&ev_end_[p]), "hipEventCreate");
        check(hipEventCreateWithFlags(&ev_done_[p], hipEventDisableTiming), "hipEventCreate");
    }
    slots_.assign(A, nullptr);
    const int b0 = batch_blocks(0), b1 = batch_blocks(1), b2 = batch_blocks(2), b3 = batch_blocks(3);
    if (b0 <= 0 || b1 <= 0 || b2 <= 0 || b3 <= 0) fail(FIU_ERR_DEVICE, "cannot size the persistent scan launches");
    // the step graph: counter reset, the four persistent scans, the merges
    char *d = desc_d_.p;
    check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "graph capture");
    const hipError_t le =
        launch_batch(jobs_d_.p, (const BatchSeg *)(d + off_segs0_), (const BatchSeg *)(d + off_segs1_),
                     (const BatchSeg *)(d + off_segs2_), (const BatchSeg *)(d + off_segs3_),
                     (const int32_t *)(d + off_merge_), (BatchHeader *)d, b0, b1, b2, b3, A, stream_);
    const hipError_t ce = hipStreamEndCapture(stream_, &graph_);
    check(le, "batched step capture");
    check(ce, "graph capture end");
    check(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0), "graph instantiate");
}

Engine::~Engine()
{
    (void)hipSetDevice(device_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    for (int p = 0; p < 2; ++p) {
        if (desc_h_[p]) (void)hipHostFree(desc_h_[p]);
        if (states_h_[p]) (void)hipHostFree(states_h_[p]);
        if (ev_start_[p]) (void)hipEventDestroy(ev_start_[p]);
        if (ev_end_[p]) (void)hipEventDestroy(ev_end_[p]);
        if (ev_done_[p]) (void)hipEventDestroy(ev_done_[p]);
    }
    if (jobs_h_) (void)hipHostFree(jobs_h_);
    if (top_h_) (void)hipHostFree(top_h_);
    if (init_h_) (void)hipHostFree(init_h_);
    if (stream_) (void)hipStreamDestroy(stream_);
    if (rb_) (void)hipStreamDestroy(rb_);
    if (prep_) (void)hipStreamDestroy(prep_);
}

void Engine::plan_job(Job *j)
{
    fiu_dataset *d = j->ds;
    const int order = j->order, M = order - 1;
    if (order < 2) fail(FIU_ERR_INVALID, "order must be >= 2");
    if (order > kMaxBatchOrder)
        fail(FIU_ERR_UNSUPPORTED, "order " + std::to_string(order) + " is above the batched scans' limit (5)");
    if (j->step < 1) fail(FIU_ERR_INVALID, "distribution step must be >= 1");
    if (j->offset < 0) fail(FIU_ERR_INVALID, "distribution offset must be >= 0");
    if (j->outputs > (uint32_t)kMaxOutputs) fail(FIU_ERR_UNSUPPORTED, "outputs above " + std::to_string(kMaxOutputs));
    const int n = (int)d->snps;
    if (binom_sat((uint64_t)n, (uint64_t)order) >= UINT64_MAX - 1)
        fail(FIU_ERR_UNSUPPORTED, "more than 2^64 - 2 combinations: their ranks do not fit the 64-bit key");
    j->n_pref = 0;
    if (n < order || j->outputs == 0) return;
    const uint64_t total_prefixes = binom_sat((uint64_t)n, (uint64_t)M);
    if (total_prefixes == UINT64_MAX || total_prefixes > (uint64_t)INT64_MAX / 2)
        fail(FIU_ERR_UNSUPPORTED, "too many prefixes");
    j->n_pref = total_prefixes > (uint64_t)j->offset
                    ? (int64_t)((total_prefixes - (uint64_t)j->offset + (uint64_t)j->step - 1) / (uint64_t)j->step)
                    : 0;
    j->rs.combinations = (j->step == 1 && j->offset == 0)
                             ? binom_sat((uint64_t)n, (uint64_t)order)
                             : count_combinations((uint64_t)n, M, (uint64_t)j->step, (uint64_t)j->offset);
    // the matrix cores take order 2 (any distribution: the tiles mask other
    // shares) and whole order-3 scans, as the single search does by default
    const int T = d->dev.n_pad / 64;
    // a tuning entry for this shape class picks the algorithm and the
    // popcount kinds' unit size
    fiu_tune_entry te{};
    const bool has_te = tuned(order, d->snps, (uint64_t)d->cases + d->ctrls, te);
    const int tuned_algo = has_te ? te.algorithm : FIU_ALGO_AUTO;
    const double unit_instr = has_te && te.batch_unit ? 40000.0 * std::ldexp(1.0, (int)te.batch_unit - 3) : 40000.0;
    if ((tuned_algo == FIU_ALGO_AUTO || tuned_algo == FIU_ALGO_MFMA) && mfma_dataset(d) && d->fp4_ready &&
        (order == 2 || (order == 3 && d->pair_ready && j->step == 1))) {
        j->kind = order == 2 ? kKindMfma2 : kKindMfma3;
        // the tile: a tuning entry's batch tile (batch_unit 1 = the 64 x 32 /
        // 64 x 64 tiles, 2 = the wide ones), else the sizes where the single
        // search takes the wide tiles (their counts wait as 16-bit pairs:
        // classes below 2^16 samples).  The entry's variant, timed on single
        // searches, is not used: on the 64-dataset batch, wide tiles for
        // the order-3 datasets of 512-2,048 SNPs that the single-search
        // tuner puts on them made the batch 9 % slower (finer units balance
        // the persistent kernel better)
        const uint64_t samples = (uint64_t)d->cases + d->ctrls;
        bool wide = samples >= 8192 && (uint64_t)n >= (order == 3 ? 2048u : 16384u);
        if (tuned_algo == FIU_ALGO_MFMA && te.batch_unit) wide = te.batch_unit == 2;
        j->wide = wide && d->cases < (1u << 16) && d->ctrls < (1u << 16);
        j->pw = 1;
        j->n_pref = order == 2 ? mfma2_superblocks(T) : std::max(0, n - 2);
        // cost of one combination in the wave-instruction currency of the
        // popcount kinds (measured: about 6x the popcount throughput)
        j->instr_per_combo = 0.15 * ((double)d->dev.quads[0] + (double)d->dev.quads[1]) + 3.0;
        j->seed = 1;
        j->rs.algorithm = FIU_ALGO_MFMA;
        return;
    }
    const bool marginal = tuned_algo != FIU_ALGO_DIRECT && d->dev.uniform &&
                          (order == 2 || (order == 3 && d->pair_ready) || order >= 4);
    // order 5 with every cell counted goes to the same launch as the
    // marginal order-5 kind (lanes split the prefix rows)
    const bool split = marginal ? order >= 4 : order == 5;
    j->kind = marginal ? (order == 2 ? kKindScan2m : order == 3 ? kKindScan3m : order == 4 ? kKindKm4 : kKindKm5)
                       : (order == 2 ? kKindDirect2 : order == 3 ? kKindDirect3 : order == 4 ? kKindDirect4
                                                                                            : kKindKm5All);
    j->pw = batch_group(j->kind);
    // cells counted per word: 2^k (marginal orders 2-3), 2 * 3^(k-1)
    // (marginal orders 4-5), 3^k (direct)
    int cells = 1;
    for (int i = 0; i < order; ++i) cells *= marginal && order <= 3 ? 2 : 3;
    if (marginal && order >= 4) cells = cells / 3 * 2;
    const double quads = (double)d->dev.quads[0] + (double)d->dev.quads[1];
    const double epi = marginal ? (order == 3 ? 600.0 : order >= 4 ? 400.0 : 200.0) : 300.0;
    const double per_tile = (double)cells * 8.0 * quads + epi;   // wave instructions per prefix and tile
    const int n_tiles = std::max(1, d->dev.n_pad / 64);
    // about 40k wave instructions (tens of microseconds) per unit, or the
    // tuned size
    j->tiles = (int)std::clamp(std::lround(unit_instr / (j->pw * per_tile)), 1L, (long)n_tiles);
    // marginal orders 4-5 count the prefix rows once per unit: at least 4
    // tiles (256 last SNPs) per unit keep that pass a few per cent
    if (split) j->tiles = std::min(n_tiles, std::max(j->tiles, 4));
    j->chunks = (n_tiles + j->tiles - 1) / j->tiles;
    const double per_pref = (double)(n - M) / (double)(M + 1);   // mean combinations per prefix
    j->instr_per_prefix = (per_pref / 64.0 + 0.5) * per_tile;
    // first slice: about 16k combinations (its exact top-N seeds the threshold)
    j->seed = std::max<int64_t>(j->pw, (int64_t)(16384.0 / std::max(1.0, per_pref)) / j->pw * j->pw);
    j->rs.algorithm = marginal ? FIU_ALGO_MARGINAL : FIU_ALGO_DIRECT;
}

// Scans the matrix cores would take (order 2, whole order-3 scans): their
// datasets get the fp4 image when it fits.
bool wants_fp4(const Job *j) { return j->order == 2 || (j->order == 3 && j->step == 1); }

void prepare_fp4(Job *j, hipStream_t s)
{
    fiu_dataset *d = j->ds;
    if (wants_fp4(j) && mfma_dataset(d) && (j->order == 2 || d->pair_ready)) (void)ensure_fp4_image(d, s);
}

// Estimated cost of units [a, b) of a job.  Order 3 on the matrix cores: first
// SNP x has C(n-1-x, 2) triples, so [a, b) has C(n-a, 3) - C(n-b, 3); order 2:
// a super-block is 64 tiles of 64 x 64 pairs.
double range_cost(const Job *j, int64_t a, int64_t b)
{
    if (b <= a) return 0.0;
    if (j->kind == kKindMfma3) {
        const double n = (double)j->ds->snps;
        auto c3 = [](double m) { return m < 3.0 ? 0.0 : m * (m - 1.0) * (m - 2.0) / 6.0; };
        return (c3(n - (double)a) - c3(n - (double)b)) * j->instr_per_combo;
    }
    if (j->kind == kKindMfma2) return (double)(b - a) * 64.0 * 4096.0 * j->instr_per_combo;
    return (double)(b - a) * j->instr_per_prefix;
}

// Largest e in (from, limit] with range_cost(from, e) <= budget (at least
// from + 1).
int64_t range_end(const Job *j, int64_t from, int64_t limit, double budget)
{
    int64_t lo = from + 1, hi = limit;
    if (lo >= hi) return hi;
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo + 1) / 2;
        if (range_cost(j, from, mid) <= budget) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

void Engine::ready(Job *j)
{
    std::lock_guard<std::mutex> lk(ready_mu_);
    ready_.push_back(j);
}

bool Engine::idle()
{
    std::lock_guard<std::mutex> lk(ready_mu_);
    return ready_.empty() && active_.empty() && step_no_ == done_no_;
}

void Engine::activate(Job *j, int slot)
{
    j->slot = slot;
    slots_[slot] = j;
    active_.push_back(j);
    if (j->staged) {
        // copies, transposes and the pair table go ahead of the job's first
        // step on the same stream; both stagings they read (pinned host,
        // device) are released once the next step, launched after them, has
        // completed
        enqueue_upload(j->ds, j->up, stream_);
        deferred_.push_back({step_no_, j->up.host, pinned_});
        deferred_.push_back({step_no_, j->up.staging_d, j->ds->arena});
        j->up.host = nullptr;
        j->up.staging_d = nullptr;
        j->staged = false;
    }
    if (slot < 0 || slot >= cfg_.max_active) fail(FIU_ERR_DEVICE, "internal: job slot out of range");
    BatchJob &B = jobs_h_[slot];
    std::memset(&B, 0, sizeof(B));
    fiu_dataset *d = j->ds;
    B.D = d->dev;
    ScanParams &P = B.P;
    P.step = j->step;
    P.offset = j->offset;
    P.total_prefixes = binom_sat(d->snps, (uint64_t)(j->order - 1));
    P.binom = j->binom;
    P.cand = cand_d_.p + (size_t)slot * cfg_.cand_cap;
    P.cand_cap = cfg_.cand_cap;
    P.order = j->order;
    P.tag = (uint32_t)slot;
    P.pair = (j->kind == kKindScan3m || j->kind == kKindMfma3) ? d->pair.p : nullptr;
    P.hist = nullptr;
    mi_constants(d->cases, d->ctrls, d->complete(), P);
    B.st = states_d_.p + slot;
    B.top = top_d_.p + (size_t)slot * kMaxOutputs;
    B.outputs = j->outputs;
    B.kind = j->kind;
    B.chunks = j->chunks;
    B.tiles = j->tiles;
    B.T = d->dev.n_pad / 64;
    B.wide = j->wide ? 1 : 0;
    if (batch_mfma(j->kind)) {
        B.F.base = d->fp4.p;
        B.F.class_off[0] = 0;
        B.F.class_off[1] = fp4_image_units(d->dev, 0);
    }
    if (!P.binom) fail(FIU_ERR_DEVICE, "internal: job without a binomial table");
    check_range(jobs_d_.p + slot, sizeof(B), jobs_d_.p, jobs_d_.n * sizeof(BatchJob), "job upload");
    check_range(states_d_.p + slot, sizeof(TopState), states_d_.p, states_d_.n * sizeof(TopState), "state init");
    check(hipMemcpyAsync(jobs_d_.p + slot, &B, sizeof(B), hipMemcpyHostToDevice, stream_), "job upload");
    check(hipMemcpyAsync(states_d_.p + slot, init_h_, sizeof(TopState), hipMemcpyHostToDevice, stream_), "state init");
}

bool Engine::launch_step()
{
    // jobs with work left
    std::vector<Job *> cand;
    for (Job *j : active_)
        if (!j->redo.empty() || j->next < j->n_pref) cand.push_back(j);
    if (cand.empty()) return false;
    const int parity = (int)(step_no_ & 1);
    StepRec &R = rec_[parity];
    R = StepRec();
    // demand per job in wave instructions, capped so that a slice is never
    // larger than everything given to the job before (doubling)
    const double budget = rate_ * cfg_.step_ms;
    std::vector<double> want(cand.size());
    for (size_t i = 0; i < cand.size(); ++i) {
        Job *j = cand[i];
        double redo = 0.0;
        int64_t redo_units = 0;
        for (auto &r : j->redo) redo += range_cost(j, r.first, r.second), redo_units += r.second - r.first;
        const int64_t cap = std::max<int64_t>(j->seed, j->given);
        want[i] = redo + range_cost(j, j->next, std::min(j->n_pref, j->next + std::max<int64_t>(0, cap - redo_units)));
    }
    // water-filling: equal shares, what a small job does not need goes to the others
    std::vector<size_t> order(cand.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return want[a] < want[b]; });
    std::vector<double> give(cand.size(), 0.0);
    double left = budget;
    for (size_t k = 0; k < order.size(); ++k) {
        const size_t i = order[k];
        const double share = left / (double)(order.size() - k);
        give[i] = std::min(want[i], share);
        left -= give[i];
    }
    char *desc = desc_h_[parity];
    BatchHeader *H = (BatchHeader *)desc;
    int32_t *merge = (int32_t *)(desc + off_merge_);
    BatchSeg *segs[3] = {(BatchSeg *)(desc + off_segs0_), (BatchSeg *)(desc + off_segs1_),
                         (BatchSeg *)(desc + off_segs3_)};
    BatchSeg *msegs = (BatchSeg *)(desc + off_segs2_);
    uint32_t nseg[3] = {0, 0, 0}, nmseg = 0;
    uint64_t units[3] = {0, 0, 0}, munits = 0;
    uint32_t nmerge = 0;
    // matrix-core ranges: per-XCD queue units (see k_batch_mfma); order-3
    // ranges are cut where the first SNP's tile changes (a constant tile
    // count per segment)
    auto add_mseg = [&](Job *j, int64_t a, int64_t b) {
        const int T = j->ds->dev.n_pad / 64;
        std::vector<std::pair<int64_t, int64_t>> parts;
        if (j->kind == kKindMfma3) {
            for (int64_t x = a; x < b;) {
                const int64_t e = std::min(b, (x + 1) / 64 * 64 + 63);   // last a with the same (a + 1) / 64, + 1
                parts.emplace_back(x, e);
                x = e;
            }
        } else {
            parts.emplace_back(a, b);
        }
        if (nmseg + parts.size() > max_msegs_) return false;
        uint64_t total = 0;
        for (auto &pr : parts) {
            uint64_t u;
            if (j->kind == kKindMfma3) {
                const int64_t m = T - (pr.first + 1) / 64, items = j->wide ? m * (m + 1) / 2 : m * (m + 1);
                u = (uint64_t)((items + 7) / 8) * (uint64_t)(pr.second - pr.first);
            } else {
                u = (uint64_t)((pr.second - pr.first + 7) / 8) * (j->wide ? 32 : 64);
            }
            total += u;
        }
        if (munits + total > (1ull << 31)) return false;
        for (auto &pr : parts) {
            uint64_t u;
            if (j->kind == kKindMfma3) {
                const int64_t m = T - (pr.first + 1) / 64, items = j->wide ? m * (m + 1) / 2 : m * (m + 1);
                u = (uint64_t)((items + 7) / 8) * (uint64_t)(pr.second - pr.first);
            } else {
                u = (uint64_t)((pr.second - pr.first + 7) / 8) * (j->wide ? 32 : 64);
            }
            BatchSeg &sg = msegs[nmseg++];
            sg.start = munits;
            sg.j0 = pr.first;
            sg.j1 = pr.second;
            sg.job = j->slot;
            sg.pad_ = 0;
            munits += u;
        }
        R.ranges.emplace_back(j, a, b);
        R.est_instr += range_cost(j, a, b);
        return true;
    };
    auto add_seg = [&](Job *j, int64_t a, int64_t b) {
        if (batch_mfma(j->kind)) return add_mseg(j, a, b);
        const int L = j->kind == kKindDirect4 ? 1
                      : (j->kind == kKindKm4 || j->kind == kKindKm5 || j->kind == kKindKm5All) ? 2
                                                                                                 : 0;
        if (nseg[L] >= max_segs_) return false;
        const uint64_t u = (uint64_t)((b - a + j->pw - 1) / j->pw) * (uint64_t)j->chunks;
        if (units[L] + u > (1ull << 31)) return false;
        BatchSeg &s = segs[L][nseg[L]++];
        s.start = units[L];
        s.j0 = a;
        s.j1 = b;
        s.job = j->slot;
        s.pad_ = 0;
        units[L] += u;
        R.ranges.emplace_back(j, a, b);
        R.est_instr += (double)(b - a) * j->instr_per_prefix;
        return true;
    };
    for (size_t i = 0; i < cand.size(); ++i) {
        Job *j = cand[i];
        double g = give[i];
        const size_t r0 = R.ranges.size();
        // re-scans first (whole ranges: they were slices before)
        bool any = false;
        while ((g > 0.0 || !any) && !j->redo.empty()) {
            const auto r = j->redo.front();
            if (!add_seg(j, r.first, r.second)) break;
            j->redo.pop_front();
            g -= range_cost(j, r.first, r.second);
            any = true;
        }
        if ((g > 0.0 || !any) && j->next < j->n_pref) {
            // at least one unit group; no more than everything given before
            const int64_t cap = j->next + std::max<int64_t>(j->seed, j->given);
            int64_t e = range_end(j, j->next, std::min(j->n_pref, std::max(cap, j->next + j->pw)), g);
            e = (e + j->pw - 1) / j->pw * j->pw;
            e = std::min(e, j->n_pref);
            if (add_seg(j, j->next, e)) {
                j->given += e - j->next;
                j->next = e;
            }
        }
        if (R.ranges.size() > r0) {
            merge[nmerge++] = j->slot;
            R.jobs.push_back(j);
            ++j->inflight;
            if (j->rs.start_ms == 0.0) j->rs.start_ms = ms_since(epoch_);
            ++j->rs.steps;
        }
    }
    if (R.jobs.empty()) return false;
    std::memset(H, 0, sizeof(BatchHeader));
    for (int L = 0; L < 3; ++L) {
        H->nseg[L] = nseg[L];
        H->units[L] = (uint32_t)units[L];
    }
    H->nmerge = nmerge;
    H->mseg = nmseg;
    H->munits = (uint32_t)munits;
    // descriptor: header + merge list + launch-0 segments in one copy, the
    // launch-1 segments in a second one when there are any; each range stays
    // inside its region of the descriptor
    if (nmerge > (uint32_t)cfg_.max_active || nseg[0] > max_segs_ || nseg[1] > max_segs_ || nseg[2] > max_segs_ ||
        nmseg > max_msegs_)
        fail(FIU_ERR_DEVICE, "internal: step descriptor overflow");
    check_range(desc + off_segs3_, nseg[2] * sizeof(BatchSeg), desc, desc_bytes_, "step descriptor");
    check(hipMemcpyAsync(desc_d_.p, desc, off_segs0_ + nseg[0] * sizeof(BatchSeg), hipMemcpyHostToDevice, stream_),
          "step descriptor");
    if (nseg[1])
        check(hipMemcpyAsync(desc_d_.p + off_segs1_, desc + off_segs1_, nseg[1] * sizeof(BatchSeg),
                             hipMemcpyHostToDevice, stream_),
              "step descriptor");
    if (nmseg)
        check(hipMemcpyAsync(desc_d_.p + off_segs2_, desc + off_segs2_, nmseg * sizeof(BatchSeg),
                             hipMemcpyHostToDevice, stream_),
              "step descriptor");
    if (nseg[2])
        check(hipMemcpyAsync(desc_d_.p + off_segs3_, desc + off_segs3_, nseg[2] * sizeof(BatchSeg),
                             hipMemcpyHostToDevice, stream_),
              "step descriptor");
    check(hipEventRecord(ev_start_[parity], stream_), "event");
    check(hipGraphLaunch(exec_, stream_), "step graph launch");
    check(hipEventRecord(ev_end_[parity], stream_), "event");
    check(hipMemcpyAsync(states_h_[parity], states_d_.p, (size_t)cfg_.max_active * sizeof(TopState),
                         hipMemcpyDeviceToHost, stream_),
          "state read");
    check(hipEventRecord(ev_done_[parity], stream_), "event");
    R.valid = true;
    ++step_no_;
    ++stats_.steps;
    stats_.units += units[0] + units[1] + units[2] + 8 * munits;
    return true;
}

void Engine::finish(Job *j, uint32_t n_top)
{
    // read into pinned memory: a direct device-to-host copy, no host-side
    // staging in the runtime
    j->count = std::min<uint32_t>({n_top, j->outputs, (uint32_t)kMaxOutputs});
    if (j->count) {
        const Cand *src = top_d_.p + (size_t)j->slot * kMaxOutputs;
        check_range(src, j->count * sizeof(Cand), top_d_.p, top_d_.n * sizeof(Cand), "top read");
        check(hipMemcpyAsync(top_h_, src, j->count * sizeof(Cand), hipMemcpyDeviceToHost, rb_), "top read");
        check(hipStreamSynchronize(rb_), "top read sync");
    }
    j->combos.assign((size_t)j->count * j->order, 0);
    j->mi.assign(j->count, 0.0f);
    unpack_results(top_h_, j->count, j->order, j->ds->snps, j->combos.data(), j->mi.data());
    slots_[j->slot] = nullptr;
    active_.erase(std::find(active_.begin(), active_.end(), j));
    j->slot = -1;
    j->rs.done_ms = ms_since(epoch_);
    j->rs.latency_ms = j->rs.done_ms - j->rs.submit_ms;
    on_done_(j);  // marks the job done (under the owner's lock)
}

void Engine::complete_step(int parity)
{
    StepRec &R = rec_[parity];
    check(hipEventSynchronize(ev_done_[parity]), "step wait");
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, ev_start_[parity], ev_end_[parity]), "event time");
    stats_.device_ms += ms;
    if (ms > 0.05f && R.est_instr > 0.0) {
        // calibrate the wave-instruction rate on steps long enough to fill the device
        const double r = R.est_instr / ms;
        if (ms > 0.5 * cfg_.step_ms) rate_ = 0.5 * rate_ + 0.5 * r;
        else rate_ = std::max(rate_, r);
    }
    const TopState *S = states_h_[parity];
    for (auto &t : R.ranges) {
        Job *j = std::get<0>(t);
        if (S[j->slot].step_overflow) {
            j->redo.emplace_back(std::get<1>(t), std::get<2>(t));
            ++stats_.redo_ranges;
        }
    }
    for (Job *j : R.jobs) --j->inflight;
    for (Job *j : R.jobs)
        if (j->slot >= 0 && j->inflight == 0 && j->next >= j->n_pref && j->redo.empty()) finish(j, S[j->slot].n_top);
    R.valid = false;
    // stagings whose copies and transposes were enqueued before this step are free
    const uint64_t finished = done_no_;
    bool freed = false;
    for (size_t i = 0; i < deferred_.size();) {
        if (deferred_[i].step <= finished) {
            if (deferred_[i].block && deferred_[i].arena) deferred_[i].arena->release(deferred_[i].block);
            deferred_[i] = deferred_.back();
            deferred_.pop_back();
            freed = true;
        } else {
            ++i;
        }
    }
    if (freed && on_staging_freed_) on_staging_freed_();
    ++done_no_;
}

bool Engine::pump()
{
    check(hipSetDevice(device_), "hipSetDevice");
    // activate ready jobs into free slots
    for (;;) {
        Job *j = nullptr;
        {
            std::lock_guard<std::mutex> lk(ready_mu_);
            if (ready_.empty()) break;
            auto it = std::find(slots_.begin(), slots_.end(), nullptr);
            if (it == slots_.end()) break;
            j = ready_.front();
            ready_.pop_front();
            if (j->n_pref > 0) activate(j, (int)(it - slots_.begin()));
        }
        if (j->n_pref == 0) {  // nothing to scan: empty result
            j->rs.start_ms = j->rs.done_ms = ms_since(epoch_);
            j->rs.latency_ms = j->rs.done_ms - j->rs.submit_ms;
            on_done_(j);
        }
    }
    const bool launched = launch_step();
    const uint64_t in_flight = step_no_ - done_no_;
    if (in_flight >= 2 || (!launched && in_flight > 0)) {
        complete_step((int)(done_no_ & 1));
        return true;
    }
    return launched;
}

// ---------------------------------------------------------------------------
// Scan server
// ---------------------------------------------------------------------------
class Server {
  public:
    Server(int device, const EngineConfig &ec, size_t budget);
    ~Server();
    uint64_t submit(const fiu_request *req);
    fiu_status wait(uint64_t id, double timeout_ms, int32_t *combos, float *mi, uint32_t *count,
                    fiu_request_stats *stats, std::string &err);
    void info(fiu_server_info *out);

  private:
    void ingest_main();
    void sched_main();
    size_t footprint(const Job *j) const;
    size_t host_footprint(const Job *j) const;
    bool fits(const Job *j) const;
    void release_staging(Job *j);

    int device_;
    Clock::time_point epoch_;
    std::unique_ptr<Arena> arena_;
    std::unique_ptr<Arena> pinned_;   // staging of host tables on their way to the device
    std::unique_ptr<Engine> eng_;
    hipStream_t ingest_stream_ = nullptr;
    std::mutex mu_;
    std::condition_variable cv_ingest_, cv_sched_, cv_done_;
    std::deque<Job *> queue_;                                 // submitted, not uploaded
    std::unordered_map<uint64_t, std::unique_ptr<Job>> jobs_;  // every job not yet collected
    uint64_t next_id_ = 1;
    uint64_t uploading_ = 0, on_device_ = 0, completed_ = 0;
    uint64_t mem_gen_ = 0;   // bumped whenever device or pinned memory is released
    bool stop_ = false;
    std::string fatal_;
    std::thread ingest_, sched_;
};

Server::Server(int device, const EngineConfig &ec, size_t budget) : device_(device), epoch_(Clock::now())
{
    check(hipSetDevice(device_), "hipSetDevice");
    if (budget == 0) {
        size_t free_b = 0, total_b = 0;
        check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        budget = std::min<size_t>(free_b / 4, 8ull << 30);
    }
    arena_ = std::make_unique<Arena>(budget, device_);
    // a staged upload holds its tables twice on the device (staging and
    // transposed image): pinned staging beyond a quarter of the budget is
    // never usable; larger tables take the synchronous path
    pinned_ = std::make_unique<Arena>(std::min<size_t>(budget / 4, 1ull << 30), device_, true);
    eng_ = std::make_unique<Engine>(device_, ec, epoch_);
    check(hipStreamCreateWithFlags(&ingest_stream_, hipStreamNonBlocking), "hipStreamCreate");
    eng_->set_staging(pinned_.get(), [this] {
        {
            std::lock_guard<std::mutex> lk(mu_);
            ++mem_gen_;
        }
        cv_ingest_.notify_all();
    });
    eng_->set_on_done([this](Job *j) {
        release_staging(j);
        j->owned.reset();  // device tables back to the arena
        std::lock_guard<std::mutex> lk(mu_);
        ++mem_gen_;
        j->done = true;
        --on_device_;
        ++completed_;
        cv_done_.notify_all();
        cv_ingest_.notify_all();
    });
    ingest_ = std::thread([this] { ingest_main(); });
    sched_ = std::thread([this] { sched_main(); });
}

Server::~Server()
{
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_ingest_.notify_all();
    cv_sched_.notify_all();
    if (ingest_.joinable()) ingest_.join();
    if (sched_.joinable()) sched_.join();
    eng_.reset();
    jobs_.clear();
    if (ingest_stream_) (void)hipStreamDestroy(ingest_stream_);
}

size_t Server::footprint(const Job *j) const
{
    if (j->ds) return 0;  // resident dataset: nothing to upload
    size_t b = dataset_bytes(j->snps, j->cw, j->tw) + staging_bytes(j->snps, j->cw, j->tw);
    b += (size_t)(j->order + 1) * (j->snps + 1) * 8 + Arena::kAlign;   // binomial table (rows up to k)
    if (j->order == 3) b += pair_table_bytes(j->snps) + Arena::kAlign;
    // the fp4 image of the matrix-core scans is not counted: it is allocated
    // at upload only when it fits, else the popcount kinds scan the dataset
    return b;
}

size_t Server::host_footprint(const Job *j) const
{
    if (j->ds) return 0;
    const size_t b = staged_host_bytes(j->snps, j->cw, j->tw, j->order);
    return b <= pinned_->capacity() ? b : 0;   // larger tables take the synchronous path
}

bool Server::fits(const Job *j) const
{
    return footprint(j) <= arena_->capacity() - arena_->used() &&
           host_footprint(j) <= pinned_->capacity() - pinned_->used();
}

void Server::release_staging(Job *j)
{
    if (!j->staged) return;
    if (j->up.host) pinned_->release(j->up.host);
    if (j->up.staging_d) arena_->release(j->up.staging_d);
    j->up = StagedUpload();
    j->staged = false;
}

uint64_t Server::submit(const fiu_request *req)
{
    auto j = std::make_unique<Job>();
    j->order = req->order;
    j->step = req->dist_step;
    j->offset = req->dist_offset;
    j->outputs = req->outputs;
    if (req->order < 2) fail(FIU_ERR_INVALID, "order must be >= 2");
    if (req->order > kMaxBatchOrder) fail(FIU_ERR_UNSUPPORTED, "order above the batched scans' limit (5)");
    if (req->dist_step < 1 || req->dist_offset < 0) fail(FIU_ERR_INVALID, "bad distribution");
    if (req->outputs > (uint32_t)kMaxOutputs) fail(FIU_ERR_UNSUPPORTED, "outputs above " + std::to_string(kMaxOutputs));
    if (req->dataset) {
        j->ds = const_cast<fiu_dataset *>(req->dataset);
        if (j->ds->device != device_) fail(FIU_ERR_INVALID, "dataset lives on another device");
    } else {
        if (!req->words && req->snps) fail(FIU_ERR_INVALID, "request has neither a dataset nor words");
        if ((uint64_t)req->cases > req->cases_words * 64 || (uint64_t)req->ctrls > req->ctrls_words * 64)
            fail(FIU_ERR_INVALID, "more samples than bits per row");
        j->snps = req->snps;
        j->cw = req->cases_words;
        j->tw = req->ctrls_words;
        j->cases = req->cases;
        j->ctrls = req->ctrls;
        if (footprint(j.get()) > arena_->capacity())
            fail(FIU_ERR_REJECTED, "request needs " + std::to_string(footprint(j.get())) +
                                       " device bytes, above the server's memory budget");
        j->words.assign(req->words, req->words + req->snps * 3 * (req->cases_words + req->ctrls_words));
    }
    std::lock_guard<std::mutex> lk(mu_);
    if (!fatal_.empty()) fail(FIU_ERR_DEVICE, "scan server stopped: " + fatal_);
    j->id = next_id_++;
    j->rs.submit_ms = ms_since(epoch_);
    Job *p = j.get();
    jobs_[p->id] = std::move(j);
    queue_.push_back(p);
    cv_ingest_.notify_all();
    return p->id;
}

void Server::ingest_main()
{
    (void)hipSetDevice(device_);
    current_device() = device_;
    uint64_t blocked_gen = ~0ull;   // memory generation at the last failed allocation
    for (;;) {
        Job *j = nullptr;
        {
            std::unique_lock<std::mutex> lk(mu_);
            // head of line: admit in arrival order when the footprint fits
            cv_ingest_.wait(lk, [&] {
                if (stop_) return true;
                if (queue_.empty()) return false;
                if (on_device_ + uploading_ == 0) return true;
                return mem_gen_ != blocked_gen && fits(queue_.front());
            });
            if (stop_) return;
            j = queue_.front();
            queue_.pop_front();
            ++uploading_;
        }
        bool ok = true, no_room = false;
        fiu_status st = FIU_OK;
        std::string msg;
        try {
            if (!j->ds && host_footprint(j) > 0) {
                // staged: host preparation now, device work queued by the scheduler
                fiu_dataset *d = stage_dataset(j->words.data(), j->snps, j->cw, j->tw, j->cases, j->ctrls, j->order,
                                               wants_fp4(j), arena_.get(), pinned_.get(), j->up);
                if (!d) {
                    no_room = true;
                } else {
                    j->owned.reset(d);
                    j->ds = d;
                    j->binom = d->binom.p;   // this job's own dataset, built for its order
                    j->staged = true;
                    std::vector<uint64_t>().swap(j->words);
                    Engine::plan_job(j);
                }
            } else {
                if (!j->ds) {
                    j->owned.reset(create_dataset(j->words.data(), j->snps, j->cw, j->tw, j->cases, j->ctrls,
                                                  arena_.get(), ingest_stream_, false));
                    j->ds = j->owned.get();
                    std::vector<uint64_t>().swap(j->words);
                }
                j->binom = ensure_binom(j->ds, j->order, ingest_stream_);
                if (j->order == 3 && j->ds->dev.uniform) (void)ensure_pair_table(j->ds, ingest_stream_);
                prepare_fp4(j, ingest_stream_);
                Engine::plan_job(j);
            }
        } catch (const Failure &e) {
            ok = false;
            st = e.status;
            msg = e.msg;
        } catch (const std::exception &e) {
            ok = false;
            st = FIU_ERR_INVALID;
            msg = e.what();
        }
        std::unique_lock<std::mutex> lk(mu_);
        --uploading_;
        if (no_room || (!ok && st == FIU_ERR_NOMEM && on_device_ > 0)) {
            // fragmented budget: wait until memory is released, then retry
            release_staging(j);
            if (!j->words.empty() || j->owned) {
                j->owned.reset();
                j->ds = nullptr;
            }
            blocked_gen = mem_gen_;
            queue_.push_front(j);
            if (on_device_ + uploading_ == 0 && no_room) {
                // nothing will free memory: give up on this request
                queue_.pop_front();
                j->status = FIU_ERR_REJECTED;
                j->error = "request does not fit in the server's memory budget";
                j->done = true;
                j->rs.done_ms = ms_since(epoch_);
                j->rs.latency_ms = j->rs.done_ms - j->rs.submit_ms;
                ++completed_;
                cv_done_.notify_all();
            }
            continue;
        }
        if (!ok) {
            release_staging(j);
            j->owned.reset();
            j->status = st == FIU_ERR_NOMEM ? FIU_ERR_REJECTED : st;
            j->error = msg;
            j->done = true;
            j->rs.done_ms = ms_since(epoch_);
            j->rs.latency_ms = j->rs.done_ms - j->rs.submit_ms;
            ++completed_;
            cv_done_.notify_all();
            continue;
        }
        j->rs.admit_ms = ms_since(epoch_);
        ++on_device_;
        eng_->ready(j);
        cv_sched_.notify_all();
    }
}

void Server::sched_main()
{
    (void)hipSetDevice(device_);
    current_device() = device_;
    try {
        for (;;) {
            if (eng_->pump()) continue;
            std::unique_lock<std::mutex> lk(mu_);
            if (stop_ && eng_->idle()) return;
            cv_sched_.wait_for(lk, std::chrono::milliseconds(2), [&] { return stop_ || !eng_->idle(); });
            if (stop_ && eng_->idle()) return;
        }
    } catch (const Failure &e) {
        std::lock_guard<std::mutex> lk(mu_);
        fatal_ = e.msg;
        for (auto &kv : jobs_)
            if (!kv.second->done) {
                kv.second->done = true;
                kv.second->status = FIU_ERR_DEVICE;
                kv.second->error = e.msg;
            }
        cv_done_.notify_all();
    }
}

fiu_status Server::wait(uint64_t id, double timeout_ms, int32_t *combos, float *mi, uint32_t *count,
                        fiu_request_stats *stats, std::string &err)
{
    std::unique_lock<std::mutex> lk(mu_);
    auto it = jobs_.find(id);
    if (it == jobs_.end()) {
        err = "unknown request id (or already collected)";
        return FIU_ERR_INVALID;
    }
    Job *j = it->second.get();
    auto ready = [&] { return j->done; };
    if (timeout_ms < 0) cv_done_.wait(lk, ready);
    else if (!cv_done_.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready)) {
        err = "request still running";
        return FIU_ERR_PENDING;
    }
    if (stats) *stats = j->rs;
    const fiu_status st = j->status;
    if (st == FIU_OK) {
        if (count) *count = j->count;
        if (combos && !j->combos.empty()) std::memcpy(combos, j->combos.data(), j->combos.size() * 4);
        if (mi && !j->mi.empty()) std::memcpy(mi, j->mi.data(), j->mi.size() * 4);
    } else {
        err = j->error;
        if (count) *count = 0;
    }
    jobs_.erase(it);
    return st;
}

void Server::info(fiu_server_info *out)
{
    std::lock_guard<std::mutex> lk(mu_);
    out->queued = queue_.size() + uploading_;
    out->active = on_device_;
    out->completed = completed_;
    out->memory_budget = arena_->capacity();
    out->memory_used = arena_->used();
    const fiu_batch_stats s = eng_->stats();
    out->steps = s.steps;
    out->device_ms = s.device_ms;
}

// step_ms = 0 takes `default_step_ms`: 24 ms for a batch of known searches
// (the drain at the end of each persistent launch costs about 0.8 ms, so
// longer steps waste less; measured 1.55 s at 6 ms, 1.19 s at 24 ms on 64
// uneven searches), 6 ms for the server, whose requests join at step
// boundaries.
EngineConfig engine_config(const fiu_batch_config *cfg, double default_step_ms)
{
    EngineConfig ec;
    ec.step_ms = default_step_ms;
    if (cfg) {
        if (cfg->struct_size != sizeof(fiu_batch_config)) fail(FIU_ERR_INVALID, "fiu_batch_config.struct_size mismatch");
        if (cfg->max_active) ec.max_active = (int)std::min<uint32_t>(cfg->max_active, 1024);
        if (cfg->step_ms > 0.0) ec.step_ms = cfg->step_ms;
        // at least twice the largest top list: an overflowing range's
        // re-scan then pushes fewer entries than the buffer holds
        if (cfg->candidate_cap) ec.cand_cap = std::max<uint32_t>(cfg->candidate_cap, 2 * kMaxOutputs);
    }
    return ec;
}

}  // namespace

struct fiu_server {
    std::unique_ptr<Server> impl;
};

extern "C" {

void fiu_batch_config_init(fiu_batch_config *cfg)
{
    if (!cfg) return;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->struct_size = sizeof(fiu_batch_config);
    cfg->max_active = 128;
    cfg->step_ms = 0.0;   // the default of each use (engine_config)
    cfg->candidate_cap = 1u << 16;
}

fiu_status fiu_search_batch(fiu_scan *scans, size_t n, const fiu_batch_config *cfg, fiu_batch_stats *stats)
{
    return guard([&] {
        if (!scans && n) fail(FIU_ERR_INVALID, "scans is NULL");
        const auto t0 = Clock::now();
        const EngineConfig ec = engine_config(cfg, 24.0);
        std::vector<std::unique_ptr<Job>> jobs;
        int device = -1;
        for (size_t i = 0; i < n; ++i) {
            fiu_scan &s = scans[i];
            s.out_count = 0;
            s.status = FIU_OK;
            if (!s.dataset) fail(FIU_ERR_INVALID, "scan " + std::to_string(i) + ": dataset is NULL");
            if (s.outputs && (!s.out_combos || !s.out_mi))
                fail(FIU_ERR_INVALID, "scan " + std::to_string(i) + ": output pointers are NULL");
            if (device < 0) device = s.dataset->device;
            if (s.dataset->device != device) fail(FIU_ERR_INVALID, "scans on more than one device");
            auto j = std::make_unique<Job>();
            j->id = i;
            j->ds = const_cast<fiu_dataset *>(s.dataset);
            j->order = s.order;
            j->step = s.dist_step;
            j->offset = s.dist_offset;
            j->outputs = s.outputs;
            j->scan = &s;
            jobs.push_back(std::move(j));
        }
        if (n == 0) {
            if (stats) *stats = fiu_batch_stats{};
            return;
        }
        Engine eng(device, ec, t0);
        size_t done = 0;
        eng.set_on_done([&](Job *j) {
            j->done = true;
            fiu_scan &s = *j->scan;
            s.status = j->status;
            s.out_count = j->count;
            if (j->count) {
                std::memcpy(s.out_combos, j->combos.data(), j->combos.size() * 4);
                std::memcpy(s.out_mi, j->mi.data(), j->mi.size() * 4);
            }
            ++done;
        });
        // table preparation (binomials, pair tables), then every scan is queued
        for (auto &jp : jobs) {
            Job *j = jp.get();
            try {
                j->binom = ensure_binom(j->ds, j->order, eng.prep_stream());
                if (j->order == 3 && j->ds->dev.uniform) (void)ensure_pair_table(j->ds, eng.prep_stream());
                prepare_fp4(j, eng.prep_stream());
                Engine::plan_job(j);
            } catch (const Failure &e) {
                j->scan->status = e.status;
                ++done;
                continue;
            }
            eng.ready(j);
        }
        while (!eng.idle()) eng.pump();
        for (auto &jp : jobs)
            if (jp->scan->status != FIU_OK) fail((fiu_status)jp->scan->status, "scan " + std::to_string(jp->id) + " failed");
        if (stats) {
            *stats = eng.stats();
            stats->total_ms = ms_since(t0);
        }
    });
}

fiu_status fiu_server_create(const fiu_batch_config *cfg, fiu_server **out)
{
    return guard([&] {
        if (!out) fail(FIU_ERR_INVALID, "out is NULL");
        const EngineConfig ec = engine_config(cfg, 6.0);
        auto s = std::make_unique<fiu_server>();
        s->impl = std::make_unique<Server>(current_device(), ec, cfg ? cfg->memory_budget : 0);
        *out = s.release();
    });
}

fiu_status fiu_server_submit(fiu_server *sv, const fiu_request *req, uint64_t *id)
{
    return guard([&] {
        if (!sv || !req || !id) fail(FIU_ERR_INVALID, "NULL argument");
        if (req->struct_size != sizeof(fiu_request)) fail(FIU_ERR_INVALID, "fiu_request.struct_size mismatch");
        *id = sv->impl->submit(req);
    });
}

fiu_status fiu_server_wait(fiu_server *sv, uint64_t id, double timeout_ms, int32_t *out_combos, float *out_mi,
                           uint32_t *out_count, fiu_request_stats *stats)
{
    if (!sv) {
        error_slot() = "server is NULL";
        return FIU_ERR_INVALID;
    }
    std::string err;
    const fiu_status st = sv->impl->wait(id, timeout_ms, out_combos, out_mi, out_count, stats, err);
    error_slot() = st == FIU_OK ? std::string() : err;
    return st;
}

fiu_status fiu_server_query(fiu_server *sv, fiu_server_info *out)
{
    return guard([&] {
        if (!sv || !out) fail(FIU_ERR_INVALID, "NULL argument");
        sv->impl->info(out);
    });
}

fiu_status fiu_server_destroy(fiu_server *sv)
{
    return guard([&] { delete sv; });
}

}  // extern "C"

"""Benchmark: exhaustive order-3 epistasis scan on one MI355X.

Workload (the headline configuration): 10,000 SNPs x 16,000 samples (8,000
cases / 8,000 controls), HWE genotypes with MAF ~ U(0.05, 0.5), seed 1, one
planted triple (17, 4242, 9001) with a random 27-cell penetrance table in
[0.3, 0.7] from seed 2; top-100 triples by mutual information.

A step is one whole order-3 scan (all 1.67e11 triples; --slices S makes a
step one Distribution(n, 2, S, i) share instead).  The metric is
genotype-combination x samples per second over the timed steps.

Also reported on the same JSON line:
  roofline      the dominant scan kernel against its peak: the order-3
                matrix-core scan (k_mfma3w: 64 x 64 (b, c) per workgroup;
                k_mfma3 for variants 3 / 4) against the dense fp4 MFMA rate,
                8 cells x 2 flops per combination x sample; the popcount scans
                against the int32 VALU peak with the reference algorithm's
                2*3^k/32 ops per combination x sample (k = 3: 1.6875);
                durations from HIP events on the library's stream around every
                scan launch;
  order2        50,000 SNPs x 32,000 samples at order 2: popcount vs int8 vs
                fp4 matrix-core scans, 128 x 128 (default), 64 x 128 and
                64 x 64 pairs per workgroup (identical results required);
  batched       64 uneven datasets, one search per call vs one batched call;
  serving       a stream of scan requests, plain vs the scan server;
  config1       the plain order-2 search of 1,000 SNPs x 4,000 samples;
  tuning        the autotuner's table (timed before the batched and served
                runs, which use it);
  cpu_baseline  the reference ThreadedSearch compiled from its own sources
                (oracle/_ref, all host cores) or, when that library is absent,
                the oracle's AVX-512 restatement, on a prefix sample of the
                same dataset.
"""
import argparse
import json
import math
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_SNPS, N_CASES, N_CTRLS = 10_000, 8_000, 8_000
PLANTED = (17, 4242, 9001)
ORDER, OUTPUTS = 3, 100
VALU_PEAK = 256 * 64 * 2.4e9          # int32 VALU lane-ops per second (MI355X)
# Bytes moved beyond L2 per triple by the order-3 scan kernels: rocprofv3
# FETCH_SIZE per dispatch x 2.0 (the factor a plain 1 GiB copy shows) +
# WRITE_SIZE x 1.0, divided by the triples of those dispatches
# (profiles/k_scan3m_order3_10000x16000.md, profiles/k_mfma3w_order3_10000x16000.md;
# k_mfma3, the 64 x 32 tile of variants 3 / 4, from its own earlier profile).
TRAFFIC_BYTES_PER_TRIPLE = {"k_scan3m": 1563.0, "k_mfma3w": 89.4, "k_mfma3": 47.7}


def alg_ops_per_combo(order, cases, ctrls):
    return 2 * 3 ** order * (math.ceil(cases / 32) + math.ceil(ctrls / 32))


def slice_combos(n, order, step, offset):
    """Exact number of combinations whose (order-1)-prefix has lex rank
    = offset (mod step); pair prefixes (order 3) in closed form."""
    assert order == 3
    total = 0
    r0 = 0
    for a in range(n - 1):
        L = n - 1 - a                      # prefixes (a, b), b = a+1 .. n-1, ranks r0 .. r0+L-1
        first = (offset - r0) % step       # first local index with rank = offset mod step
        if first < L:
            cnt = (L - 1 - first) // step + 1
            # b = a+1+t, combos (n-1-b) = L-1-t, for t = first, first+step, ...
            t_last = first + (cnt - 1) * step
            total += cnt * (L - 1) - (first + t_last) * cnt // 2
        r0 += L
    return total


def threaded_combos(n, order, step, offset, threads):
    """Combinations ThreadedSearch(threads) evaluates on Distribution(n,
    order-1, step, offset): thread t scans layer(threads, t) = (step *
    threads, offset * threads + t) (ThreadedSearch.h:219-239,
    Distribution.h:172-176)."""
    return sum(slice_combos(n, order, step * threads, offset * threads + t) for t in range(threads))


def cpu_baseline(words, raw_path, cw, tw, threads, sample_step):
    """Reference ThreadedSearch (or the oracle port, which splits the work
    the same way) on Distribution(n, 2, sample_step, 0); with the reference
    build, also the reference loader's time on the .raw."""
    import oracle
    combos = threaded_combos(N_SNPS, ORDER, sample_step, 0, threads)
    ref = oracle.ref("avx512")
    read_s = None
    if ref is not None and oracle.has_avx512():
        t0 = time.perf_counter()
        d = ref.load(raw_path)
        read_s = time.perf_counter() - t0
        _, _, sec = d.search(ORDER, OUTPUTS, step=sample_step, offset=0, threads=threads)
        kind = "reference"
    else:
        d = oracle.Dataset.__new__(oracle.Dataset)
        ds = oracle._DS(N_SNPS, N_CASES, N_CTRLS, cw, tw, words.ctypes.data_as(oracle.C.POINTER(oracle.C.c_uint64)))
        t0 = time.perf_counter()
        oracle.lib().orc_search(oracle.C.byref(ds), ORDER, sample_step, 0, OUTPUTS, threads, 0,
                                np.zeros(OUTPUTS * ORDER, np.int32), np.zeros(OUTPUTS, np.float32), None)
        sec = time.perf_counter() - t0
        kind = "port"
    return {"value": combos * (N_CASES + N_CTRLS) / sec, "unit": "combo*samples/s", "cores": threads,
            "kind": kind, "reference_read_s": read_s,
            "sample": f"Distribution({N_SNPS}, {ORDER - 1}, {sample_step}, 0) of the same dataset "
                      f"(read from a PLINK .raw by the reference loader), whose {threads} threads scan its "
                      f"layers ({sample_step * threads}, t), t < {threads}: "
                      f"{combos} of {math.comb(N_SNPS, ORDER)} triples, {sec:.1f} s, "
                      f"ThreadedSearch({threads}) default AVX-512 build" if kind == "reference" else
                      f"Distribution({N_SNPS}, {ORDER - 1}, {sample_step}, 0), layers ({sample_step * threads}, t), "
                      f"t < {threads}: {combos} triples, {sec:.1f} s, oracle AVX-512 restatement, {threads} threads"}


def mi_error(ds, geno, combos, mi):
    """MI of the headline top list against the reference's own FP32 value
    (MutualInformation<float>::compute of the same counts, from the reference
    build or the oracle's bit-exact restatement of it) and against FP64; the
    GPU's counts of those combinations against a count of the genotypes."""
    import fiuncho_amd as fa
    import oracle
    counts, _ = fa.contingency_batch(ds, combos)
    is_case = np.arange(N_CASES + N_CTRLS) < N_CASES
    ref = oracle.ref("avx512") if oracle.has_avx512() else None
    rel, ab, exact = [], [], True
    for i, c in enumerate(combos):
        cell = geno[c[0]].astype(np.int64) * 9 + geno[c[1]] * 3 + geno[c[2]]
        cc = np.bincount(cell[is_case], minlength=27)
        ct = np.bincount(cell[~is_case], minlength=27)
        exact &= bool(np.array_equal(counts[i, 0, :27], cc) and np.array_equal(counts[i, 1, :27], ct))
        n = float(N_CASES + N_CTRLS)
        pc, pt = cc / n, ct / n
        px = pc + pt

        def h(p):
            p = p[p > 0]
            return float(-(p * np.log(p)).sum())

        m64 = h(px) + h(np.array([N_CASES / n, N_CTRLS / n])) - h(np.r_[pc, pt])
        r = ref.mi(cc, ct, 3, N_CASES, N_CTRLS) if ref is not None else \
            oracle.mi_f32_ref512(cc, ct, 3, N_CASES, N_CTRLS)
        rel.append(abs(float(mi[i]) - r) / abs(r))
        ab.append(abs(float(mi[i]) - m64))
    return {"results": len(combos), "counts_exact": exact,
            "vs_reference_fp32": {"max_rel": max(rel), "median_rel": float(np.median(rel)),
                                  "source": "reference build (oracle/_ref)" if ref is not None else
                                            "oracle restatement of the reference's FP32 order"},
            "vs_fp64": {"max_abs": max(ab), "median_abs": float(np.median(ab))}}


# ---------------------------------------------------------------------------
# Batched and served forms against the plain one-search-per-call form
# ---------------------------------------------------------------------------
def uneven_specs(count, seed0, scale):
    """Uneven scans: samples ~ U{1,000..50,000}, case fraction ~ U(0.3, 0.7),
    order 2 or 3 (every 16th dataset order 4), MAF ~ U(0.05, 0.5), one seed per
    dataset (seed0 + i).  SNP counts per order come from `scale`."""
    out = []
    for i in range(count):
        r = np.random.default_rng(seed0 + i)
        order = 4 if i % 16 == 15 else int(r.choice([2, 3]))
        lo, hi, nlo, nhi = scale[order]
        samples = int(r.integers(nlo, nhi + 1))
        frac = float(r.uniform(0.3, 0.7))
        cases = max(1, int(round(samples * frac)))
        out.append(dict(order=order, snps=int(r.integers(lo, hi + 1)), cases=cases, ctrls=samples - cases,
                        seed=seed0 + i))
    return out


# SNP and sample ranges per order: the uneven-batch workload keeps the stated
# order-2 shapes (500..20,000 SNPs, 1,000..50,000 samples) and scales the
# order-3 / order-4 SNP counts so the whole batch is seconds, not hours.
BATCH_SCALE = {2: (500, 20000, 1000, 50000), 3: (200, 1500, 1000, 50000), 4: (100, 300, 1000, 5000)}
# Served requests: small scans, about 1-20 ms each.
SERVE_SCALE = {2: (500, 4000, 1000, 20000), 3: (100, 600, 1000, 20000), 4: (30, 80, 1000, 5000)}
BATCH_OUTPUTS = 50


def make_words(sp):
    import fiuncho_amd as fa
    spec = fa.synth_spec(sp["snps"], sp["cases"], sp["ctrls"], seed=sp["seed"])
    words, _, cw, tw = fa.synth_generate(spec, words=True)
    return words, cw, tw


def combo_samples(sp):
    return math.comb(sp["snps"], sp["order"]) * (sp["cases"] + sp["ctrls"])


def bench_batch(count=64):
    """64 uneven datasets (seeds 100..163), top-50 each: every search one call
    (plain) against all of them in one fiu_search_batch (batched).  Datasets
    are resident and their tables prepared before either is timed; each form
    runs twice and its faster pass is reported."""
    import fiuncho_amd as fa
    specs = uneven_specs(count, 100, BATCH_SCALE)
    dss = []
    for sp in specs:
        words, cw, tw = make_words(sp)
        d = fa.Dataset.from_words(words, sp["snps"], cw, tw, sp["cases"], sp["ctrls"])
        del words
        # binomial and pair tables (a one-prefix search builds them)
        fa.GpuSearch().run_arrays(d, sp["order"], fa.Distribution(sp["snps"], sp["order"] - 1, 1 << 30, 0), 1)
        dss.append(d)
    print(f"[bench] batch: {count} datasets resident", file=sys.stderr, flush=True)
    scans = [(d, sp["order"], fa.Distribution(sp["snps"], sp["order"] - 1), BATCH_OUTPUTS) for d, sp in zip(dss, specs)]
    # two passes of each form, alternating; the faster of each is kept (the
    # first batched call also captures its step graph)
    passes = {"plain": [], "batched": []}
    same = True
    for _ in range(2):
        t0 = time.perf_counter()
        plain = [fa.GpuSearch().run_arrays(*s) for s in scans]
        passes["plain"].append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        batched, st = fa.search_batch(scans)
        passes["batched"].append(time.perf_counter() - t0)
        print(f"[bench] batch: plain {passes['plain'][-1]:.2f} s, batched {passes['batched'][-1]:.2f} s",
              file=sys.stderr, flush=True)
        same = same and all(a[0].tolist() == b[0].tolist() and np.array_equal(a[1], b[1])
                            for a, b in zip(plain, batched))
    t_plain, t_batch = min(passes["plain"]), min(passes["batched"])
    work = sum(combo_samples(sp) for sp in specs)
    orders = [sp["order"] for sp in specs]
    for d in dss:
        d.close()
    return {"workload": f"{count} uneven datasets (seeds 100..{99 + count}): order 2 with 500..20,000 SNPs, "
                        "order 3 with 200..1,500 SNPs, order 4 (every 16th) with 100..300 SNPs and "
                        "1,000..5,000 samples; 1,000..50,000 samples, case fraction 0.3..0.7, top-50",
            "orders": {str(k): orders.count(k) for k in (2, 3, 4)},
            "combo_samples": work, "plain_s": t_plain, "batched_s": t_batch, "passes_s": passes,
            "speedup": t_plain / t_batch, "batched_combo_samples_per_s": work / t_batch,
            "steps": st.steps, "device_s": st.device_ms / 1e3, "results_identical": same}


def bench_tune():
    """The autotuner over its default shape classes (orders 2 and 3), then
    the batched popcount kernels' unit sizes over its batch classes (orders
    4 and 5); the table it loads is used by the searches that follow
    (batched, served)."""
    from fiuncho_amd import tune
    t0 = time.perf_counter()
    tune.autotune()
    dt = time.perf_counter() - t0
    t1 = time.perf_counter()
    entries = tune.autotune_batch()
    dtb = time.perf_counter() - t1
    print(f"[bench] tune: {len(entries)} classes in {dt:.1f} s, batch units in {dtb:.1f} s", file=sys.stderr,
          flush=True)
    names = {1: "direct", 2: "marginal", 3: "mfma"}
    return {"seconds": dt, "batch_seconds": dtb, "classes": [
        {"order": e.order, "snps": [e.snps_lo, e.snps_hi], "samples": [e.samples_lo, e.samples_hi],
         "algorithm": names.get(e.algorithm, e.algorithm), "variant": e.variant, "ms": e.ms,
         "batch_unit": e.batch_unit} for e in entries]}


def bench_config1(repeats=20):
    """The plain order-2 parity configuration: 1,000 SNPs x 4,000 samples
    (2,000 / 2,000), top-10, one search per call."""
    import fiuncho_amd as fa
    spec = fa.synth_spec(1000, 2000, 2000, seed=1)
    d = fa.Dataset.synth(spec)
    s = fa.GpuSearch()
    dist = fa.Distribution(1000, 1)
    s.run_arrays(d, 2, dist, 10)
    ts = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        s.run_arrays(d, 2, dist, 10)
        ts.append(time.perf_counter() - t0)
    d.close()
    ts.sort()
    med = ts[len(ts) // 2]
    print(f"[bench] config 1: {med * 1e3:.2f} ms per search (median of {repeats})", file=sys.stderr, flush=True)
    return {"workload": "order-2 exhaustive scan, 1,000 SNPs x 4,000 samples (2,000/2,000), top-10",
            "median_ms": med * 1e3, "min_ms": ts[0] * 1e3,
            "combo_samples_per_s": 499500 * 4000 / med, "algorithm": s.last_stats.algorithm}


def bench_high_orders():
    """Orders 4 to 6 (the last-SNP-marginal popcount scan, k_scan_km) on
    seeded HWE datasets of 16,000 samples (8,000 / 8,000), top-100; order 4
    also with the all-cells direct kernel, the reference's algorithm; orders
    7 to 9 with the all-cells per-workgroup kernel (k_scan_any)."""
    import fiuncho_amd as fa
    out = {}
    for order, snps in ((4, 400), (5, 100), (6, 40), (7, 20), (8, 16), (9, 14)):
        spec = fa.synth_spec(snps, 8000, 8000, seed=5)
        d = fa.Dataset.synth(spec)
        dist = fa.Distribution(snps, order - 1)
        res = {}
        for name, algo in (("marginal", fa.ALGO_MARGINAL), ("direct", fa.ALGO_DIRECT)):
            if (order in (5, 6) and algo == fa.ALGO_DIRECT) or (order >= 7 and algo == fa.ALGO_MARGINAL):
                continue
            s = fa.GpuSearch(algorithm=algo)
            s.run_arrays(d, order, fa.Distribution(snps, order - 1, 64, 63), 100)   # code and tables
            t0 = time.perf_counter()
            c, m = s.run_arrays(d, order, dist, 100)
            wall = time.perf_counter() - t0
            st = s.last_stats
            cs = st.combinations * 16000 / (st.scan_ms / 1e3)
            # the reference's AND + popcount ops (3^k cells x 2 per 32 samples)
            ref_ops = cs * 2 * 3 ** order / 32
            res[name] = {"wall_s": wall, "scan_ms": st.scan_ms, "combinations": st.combinations,
                         "combo_samples_per_s": cs, "reference_algorithm_valu_frac": ref_ops / VALU_PEAK}
            if name == "marginal":
                # executed: 2 * 3^(k-1) cells x (AND + popcount) per 32 samples
                res[name]["executed_valu_frac"] = cs * 2 * 2 * 3 ** (order - 1) / 32 / VALU_PEAK
            res[name + "_result"] = (c, m)
            print(f"[bench] order {order} {snps} x 16000 {name}: {st.scan_ms:.1f} ms scan, {cs:.3e} combo*samples/s",
                  file=sys.stderr, flush=True)
        if "direct" in res and "marginal" in res:
            res["identical"] = bool(res["marginal_result"][0].tolist() == res["direct_result"][0].tolist() and
                                    np.array_equal(res["marginal_result"][1], res["direct_result"][1]))
            res["speedup"] = res["direct"]["scan_ms"] / res["marginal"]["scan_ms"]
        res.pop("marginal_result", None)
        res.pop("direct_result", None)
        res["workload"] = f"order-{order} exhaustive scan, {snps} SNPs x 16,000 samples (8,000/8,000), top-100"
        out[f"order{order}"] = res
        d.close()
    return out


MFMA_I8_PEAK = 5.0e15    # dense int8 MFMA ops per second (MI355X)
MFMA_FP4_PEAK = 10.1e15  # dense fp4 MFMA flops per second (MI355X)


def bench_order2(snps=50_000, half=16_000, outputs=1000):
    """Order-2 exhaustive scan, 50,000 SNPs x 32,000 samples (16,000 /
    16,000), MAF ~ U(0.05, 0.5) with 20% of SNPs ~ U(0.01, 0.05), top-1,000:
    the popcount scan against the matrix-core scans (int8 operands; fp4
    operands with 128 x 128 pairs per workgroup, the default, with 64 x 128
    and with 64 x 64), same dataset, identical results required."""
    import fiuncho_amd as fa
    spec = fa.synth_spec(snps, half, half, seed=1, low_frac=0.2)
    d = fa.Dataset.synth(spec)
    dist = fa.Distribution(snps, 1)
    out = {}
    res = {}
    for name, algo, variant in (("popcount", fa.ALGO_MARGINAL, 0), ("mfma_int8", fa.ALGO_MFMA, 1),
                                ("mfma_64x64", fa.ALGO_MFMA, 2), ("mfma_64x128", fa.ALGO_MFMA, 3),
                                ("mfma", fa.ALGO_MFMA, 0)):
        s = fa.GpuSearch(algorithm=algo, variant=variant)
        s.run_arrays(d, 2, dist, outputs)   # warm
        best = None
        for _ in range(3):
            t0 = time.perf_counter()
            c, m = s.run_arrays(d, 2, dist, outputs)
            el = time.perf_counter() - t0
            if best is None or el < best[0]:
                best = (el, s.last_stats.scan_ms, s.last_stats.launches)
        res[name] = (c, m)
        out[name] = {"wall_s": best[0], "scan_ms": best[1], "launches": best[2],
                     "combo_samples_per_s": math.comb(snps, 2) * 2 * half / best[0]}
        print(f"[bench] order 2 {name}: {best[0]:.3f} s wall, {best[1]:.1f} ms scan", file=sys.stderr, flush=True)
    d.close()
    # matrix-core work actually issued: 4 products per pair and class over
    # the padded samples (128-sample quads), int8 MACs = 2 ops
    quads = math.ceil(math.ceil(half / 512) * 8 / 2)
    tiles = (math.ceil(snps / 64) * (math.ceil(snps / 64) + 1)) // 2
    mfma_ops = tiles * 64 * 64 * 4 * 2 * (quads * 128) * 2
    out["mfma_int8"]["mfma_tops"] = mfma_ops / (out["mfma_int8"]["scan_ms"] / 1e3) / 1e12
    out["mfma_int8"]["frac_of_peak"] = mfma_ops / (out["mfma_int8"]["scan_ms"] / 1e3) / MFMA_I8_PEAK
    out["mfma_int8"]["peak"] = "int8 dense 5.0 POP/s"
    for k, tile in (("mfma_64x64", "64 x 64"), ("mfma_64x128", "64 x 128"), ("mfma", "128 x 128")):
        out[k]["operands"] = "fp4 (e2m1) x fp4, fp32 accumulation (exact: counts < 2^24)"
        out[k]["pairs_per_workgroup"] = tile
        out[k]["mfma_tops"] = mfma_ops / (out[k]["scan_ms"] / 1e3) / 1e12
        out[k]["frac_of_peak"] = mfma_ops / (out[k]["scan_ms"] / 1e3) / MFMA_FP4_PEAK
        out[k]["peak"] = "fp4 dense 10.1 PFLOP/s"
    same = all(res["popcount"][0].tolist() == res[k][0].tolist() and np.array_equal(res["popcount"][1], res[k][1])
               for k in ("mfma_int8", "mfma_64x64", "mfma_64x128", "mfma"))
    return {"workload": f"order-2 exhaustive scan, {snps:,} SNPs x {2 * half:,} samples ({half:,}/{half:,}), "
                        "20% of SNPs with MAF < 0.05, top-1,000",
            **out, "int8_kernel_speedup": out["popcount"]["scan_ms"] / out["mfma_int8"]["scan_ms"],
            "kernel_speedup": out["popcount"]["scan_ms"] / out["mfma"]["scan_ms"],
            "wall_speedup": out["popcount"]["wall_s"] / out["mfma"]["wall_s"], "results_identical": same,
            "kept": "mfma (AUTO default for order 2)" if out["mfma"]["scan_ms"] < out["popcount"]["scan_ms"] else
                    "popcount"}


def percentile(xs, q):
    return float(np.percentile(np.asarray(xs, float), q)) if xs else None


def bench_serve(count=3000, distinct=300, load=0.7):
    """Scan-request stream: `count` requests over `distinct` tables at
    serving scale (table i: seed 1000 + i; request r uses table r mod
    `distinct`), each request carrying its host tables.  Capacity of each
    form with every request queued at once, then a Poisson stream at `load`
    x the served capacity through both forms (plain: one worker, one request
    at a time, FIFO).  Latency = completion - arrival."""
    import fiuncho_amd as fa
    specs = uneven_specs(distinct, 1000, SERVE_SCALE)
    tables = []
    for sp in specs:
        words, cw, tw = make_words(sp)
        tables.append(dict(order=sp["order"], outputs=BATCH_OUTPUTS, words=words, snps=sp["snps"],
                           cases_words=cw, ctrls_words=tw, cases=sp["cases"], ctrls=sp["ctrls"]))
    stream = [i % distinct for i in range(count)]
    work = sum(combo_samples(specs[i]) for i in stream)
    print(f"[bench] serve: {distinct} tables generated, {count} requests", file=sys.stderr, flush=True)

    def plain_one(r):
        d = fa.Dataset.from_words(r["words"], r["snps"], r["cases_words"], r["ctrls_words"], r["cases"], r["ctrls"])
        out = fa.GpuSearch().run_arrays(d, r["order"], fa.Distribution(r["snps"], r["order"] - 1), r["outputs"])
        d.close()
        return out

    def same_as_plain(res):
        return all(a[0].tolist() == plain_res[i][0].tolist() and np.array_equal(a[1], plain_res[i][1])
                   for i, a in zip(stream, res))

    # capacity, every request queued at t = 0
    t0 = time.perf_counter()
    plain_res = {}
    for i in stream:
        out = plain_one(tables[i])
        plain_res.setdefault(i, out)
    t_plain = time.perf_counter() - t0
    print(f"[bench] serve: plain capacity run {t_plain:.2f} s", file=sys.stderr, flush=True)
    srv = fa.Server()
    t0 = time.perf_counter()
    ids = [srv.submit(**tables[i]) for i in stream]
    served_res = [srv.wait(x) for x in ids]
    t_served = time.perf_counter() - t0
    same = same_as_plain(served_res)
    del served_res
    print(f"[bench] serve: capacity plain {count / t_plain:.1f} req/s, served {count / t_served:.1f} req/s",
          file=sys.stderr, flush=True)
    # Poisson arrivals at `load` x the served capacity
    rate = load * count / t_served
    gaps = np.random.default_rng(7).exponential(1.0 / rate, size=count)
    arrivals = np.cumsum(gaps)

    t0 = time.perf_counter()
    sub = []
    for i, a in zip(stream, arrivals):
        dt = a - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        sub.append(srv.submit(**tables[i]))
    stream_res = [srv.wait(x) for x in sub]
    lat_served = [x[2].latency_ms for x in stream_res]
    t_stream_served = time.perf_counter() - t0
    same = same and same_as_plain(stream_res)
    del stream_res
    srv.close()
    print(f"[bench] serve: served stream p50 {percentile(lat_served, 50):.1f} ms, "
          f"p99 {percentile(lat_served, 99):.1f} ms", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    lat_plain = []
    for i, a in zip(stream, arrivals):
        dt = a - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        plain_one(tables[i])
        lat_plain.append((time.perf_counter() - t0 - a) * 1e3)
    t_stream_plain = time.perf_counter() - t0
    print(f"[bench] serve: plain stream p50 {percentile(lat_plain, 50):.1f} ms, "
          f"p99 {percentile(lat_plain, 99):.1f} ms", file=sys.stderr, flush=True)
    orders = [specs[i]["order"] for i in stream]
    return {"workload": f"{count} requests over {distinct} tables (table i: seed 1000 + i, request r uses table "
                        f"r mod {distinct}): order 2 with 500..4,000 SNPs, order 3 with 100..600 SNPs, order 4 "
                        "(every 16th table) with 30..80 SNPs; 1,000..20,000 samples (order 4: ..5,000), case "
                        "fraction 0.3..0.7, top-50; host tables uploaded per request",
            "requests": count, "distinct_tables": distinct,
            "orders": {str(k): orders.count(k) for k in (2, 3, 4)},
            "combo_samples": work,
            "capacity_req_per_s": {"plain": count / t_plain, "served": count / t_served},
            "throughput_speedup": t_plain / t_served,
            "served_combo_samples_per_s": work / t_served,
            "stream": {"arrival": "Poisson", "offered_req_per_s": rate, "load_vs_served_capacity": load,
                       "served": {"p50_ms": percentile(lat_served, 50), "p99_ms": percentile(lat_served, 99),
                                  "wall_s": t_stream_served},
                       "plain": {"p50_ms": percentile(lat_plain, 50), "p99_ms": percentile(lat_plain, 99),
                                 "wall_s": t_stream_plain}},
            "results_identical": same}


def bench_ingest(path, geno, words):
    """Start-up: the headline dataset's PLINK .raw (10,000 SNPs x 16,000
    samples) read by the library -- host parse, device pack -- against host
    parse + host pack, and the device pack alone from in-memory genotypes."""
    import fiuncho_amd as fa
    out = {"file": f"PLINK .raw, {N_SNPS} SNPs x {N_CASES + N_CTRLS} samples", "file_bytes": os.path.getsize(path)}
    t0 = time.perf_counter()
    d = fa.Dataset.read([path])
    out["read_device_pack_s"] = time.perf_counter() - t0
    dev_words = d.words()
    d.close()
    t0 = time.perf_counter()
    host_words = fa.read_files([path])[0]
    out["read_host_pack_s"] = time.perf_counter() - t0
    # the device-packed tables equal the host-packed ones; both differ from
    # the generator's words only in the reference readers' re-flushed word
    # (the 8,000th case is followed by a control: case word 124 is zero)
    out["identical_words"] = bool(np.array_equal(dev_words, host_words))
    out["differs_from_generated_words"] = int(np.count_nonzero(dev_words != words))
    del dev_words, host_words
    is_ctrl = np.r_[np.zeros(N_CASES, np.uint8), np.ones(N_CTRLS, np.uint8)]
    ts = []
    for _ in range(4):
        t0 = time.perf_counter()
        fa.Dataset.from_genotypes(geno, is_ctrl).close()
        ts.append(time.perf_counter() - t0)
    out["from_genotypes_ms"] = min(ts[1:]) * 1e3
    out["genotype_bytes"] = int(geno.nbytes)
    print(f"[bench] ingest: read {out['read_device_pack_s']:.2f} s (device pack) vs "
          f"{out['read_host_pack_s']:.2f} s (host pack); from genotypes {out['from_genotypes_ms']:.1f} ms",
          file=sys.stderr, flush=True)
    return out


def write_raw(path, geno, cases):
    """PLINK .raw (FID IID PAT MAT SEX PHENOTYPE + one column per SNP)."""
    snps, n = geno.shape
    digits = np.array([ord("0"), ord("1"), ord("2")], np.uint8)
    with open(path, "wb") as f:
        f.write(("FID IID PAT MAT SEX PHENOTYPE " + " ".join(f"S{i}_A" for i in range(snps)) + "\n").encode())
        block = 512
        for j0 in range(0, n, block):
            j1 = min(n, j0 + block)
            g = digits[geno[:, j0:j1].T]                       # samples x snps
            body = np.full((j1 - j0, 2 * snps), ord(" "), np.uint8)
            body[:, 1::2] = g
            lines = []
            for j in range(j0, j1):
                ph = 2 if j < cases else 1
                lines.append(f"i{j} i{j} 0 0 0 {ph}".encode() + body[j - j0].tobytes() + b"\n")
            f.write(b"".join(lines))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--slices", type=int, default=1,
                    help="a step scans Distribution(n, 2, SLICES, i): 1 = the whole order-3 scan")
    ap.add_argument("--algorithm", default="auto", choices=["auto", "direct", "marginal", "mfma"])
    ap.add_argument("--variant", type=int, default=0, help="kernel variant of the algorithm (0 = default)")
    ap.add_argument("--cpu-sample-step", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the uneven-batch comparison")
    ap.add_argument("--no-serve", action="store_true", help="skip the served-stream comparison")
    ap.add_argument("--no-order2", action="store_true", help="skip the order-2 popcount / matrix-core comparison")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (batch / serve use the defaults)")
    ap.add_argument("--no-ingest", action="store_true", help="skip the .raw read / device packing timing")
    ap.add_argument("--no-high", action="store_true", help="skip the order-4 to order-6 scans")
    ap.add_argument("--batch-datasets", type=int, default=64)
    ap.add_argument("--serve-requests", type=int, default=3000)
    ap.add_argument("--serve-tables", type=int, default=300)
    ap.add_argument("--only-extra", action="store_true",
                    help="run only the batch / serve comparisons (no headline scan)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run one device-to-device copy of BYTES (a plain copy kernel of known "
                         "size, used to scale the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        n = args.copy_probe // 4
        src = torch.ones(n, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dst.copy_(src)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        print(json.dumps({"copy_bytes": 4 * n, "seconds": dt, "GBps": 4 * n / dt / 1e9}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU (a single MI355X); --gpus above 1 is not supported. "
              "The scan shards by Distribution layers, one search per GPU, with a host merge "
              "of the per-GPU top-N lists.", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1 or args.slices < 1 or args.warmup < 0 or args.warmup > 4096:
        sys.exit("--steps and --slices must be >= 1, --warmup within 0..4096")

    import fiuncho_amd as fa

    if args.only_extra:
        extra = {}
        if not args.no_order2:
            extra["config1"] = bench_config1()
            extra["order2"] = bench_order2()
        if not args.no_high:
            extra["high_orders"] = bench_high_orders()
        if not args.no_tune and not (args.no_batch and args.no_serve):
            extra["tuning"] = bench_tune()
        if not args.no_batch:
            extra["batched"] = bench_batch(args.batch_datasets)
        if not args.no_serve:
            extra["serving"] = bench_serve(args.serve_requests, min(args.serve_tables, args.serve_requests))
        if not args.no_ingest:
            spec = fa.synth_spec(N_SNPS, N_CASES, N_CTRLS, seed=1, planted=PLANTED, penetrance_seed=2)
            words, geno, _, _ = fa.synth_generate(spec, words=True, geno=True)
            with tempfile.TemporaryDirectory(dir="/dev/shm" if os.path.isdir("/dev/shm") else None) as tmp:
                path = os.path.join(tmp, "bench.raw")
                write_raw(path, geno, N_CASES)
                extra["ingest"] = bench_ingest(path, geno, words)
        print(json.dumps(extra), flush=True)
        return
    algo = {"auto": fa.ALGO_AUTO, "direct": fa.ALGO_DIRECT, "marginal": fa.ALGO_MARGINAL,
            "mfma": fa.ALGO_MFMA}[args.algorithm]
    t0 = time.perf_counter()
    spec = fa.synth_spec(N_SNPS, N_CASES, N_CTRLS, seed=1, planted=PLANTED, penetrance_seed=2)
    words, geno, cw, tw = fa.synth_generate(spec, words=True, geno=not (args.no_cpu and args.no_ingest))
    ds = fa.Dataset.from_words(words, N_SNPS, cw, tw, N_CASES, N_CTRLS)
    print(f"[bench] dataset generated and uploaded in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    search = fa.GpuSearch(algorithm=algo, variant=args.variant)
    S = args.slices
    # warmup (never timed): small shares of the scan build the pair table and
    # load the code; then the timed steps
    for w in range(args.warmup):
        search.run_arrays(ds, ORDER, fa.Distribution(N_SNPS, 2, 4096, 4095 - w), OUTPUTS)
    results = {}   # combination -> MI (whole-scan steps repeat the same list)
    combos = 0
    scan_ms = 0.0
    launches = 0
    used_algo = None
    t_start = time.perf_counter()
    for i in range(args.steps):
        c, m = search.run_arrays(ds, ORDER, fa.Distribution(N_SNPS, 2, S, i % S), OUTPUTS)
        st = search.last_stats
        scan_ms += st.scan_ms
        launches += st.launches
        used_algo = st.algorithm
        combos += st.combinations
        results.update(zip(map(tuple, c.tolist()), m.tolist()))
        print(f"[bench] step {i} done: {st.scan_ms / 1e3:.2f} s device", file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t_start
    top = sorted(results.items(), key=lambda r: (-r[1], r[0]))[:OUTPUTS]
    unit_samples = N_CASES + N_CTRLS
    value = combos * unit_samples / elapsed
    algo_name = {1: "direct", 2: "marginal", 3: "mfma"}.get(used_algo, str(used_algo))
    kernel = {"mfma": "k_mfma3" if search.last_stats.variant in (3, 4) else "k_mfma3w",
              "marginal": "k_scan3m"}.get(algo_name, "k_scan_direct<3>")
    per_triple = TRAFFIC_BYTES_PER_TRIPLE.get(kernel)
    traffic = per_triple * combos / max(launches, 1) if per_triple else None
    ref_ops = combos * alg_ops_per_combo(ORDER, N_CASES, N_CTRLS)   # the reference's AND + popcount ops
    if algo_name == "mfma":
        # matrix-core work: 8 cells x (cases + ctrls) fp4 multiply-adds per triple
        flops = combos * 8 * 2 * unit_samples
        roofline = {"bound": "mfma", "achieved": flops / (scan_ms / 1e3) / 1e12, "peak": MFMA_FP4_PEAK / 1e12,
                    "unit": "TFLOP/s", "frac": flops / (scan_ms / 1e3) / MFMA_FP4_PEAK, "traffic": traffic,
                    "kernel": kernel, "operands": "fp4 (e2m1) x fp4, fp32 accumulation (exact)",
                    "flops_per_combo_sample": 16,
                    "reference_algorithm_valu_frac": ref_ops / (scan_ms / 1e3) / VALU_PEAK}
    else:
        # operations the kernel executes per triple: AND + popcount of the 8
        # (marginal) or 27 (direct) cells over every stored 32-bit word
        executed = combos * 2 * (8 if algo_name == "marginal" else 27) * 2 * (cw + tw)
        roofline = {"bound": "valu", "achieved": ref_ops / (scan_ms / 1e3) / 1e12, "peak": VALU_PEAK / 1e12,
                    "unit": "TOP/s", "frac": ref_ops / (scan_ms / 1e3) / VALU_PEAK, "traffic": traffic,
                    "kernel": kernel,
                    "ops_per_combo_sample": alg_ops_per_combo(ORDER, N_CASES, N_CTRLS) / unit_samples,
                    "executed_frac": executed / (scan_ms / 1e3) / VALU_PEAK}
    roofline.update({"launches": launches, "avg_launch_ms": scan_ms / max(launches, 1)})
    line = {
        "metric": "genotype-combos*samples/s at order 3",
        "value": value,
        "unit": "combo*samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded HWE genotypes, planted triple 17/4242/9001)",
        "config": {"workload": "order-3 exhaustive scan, 10,000 SNPs x 16,000 samples (8,000/8,000), top-100",
                   "step": "one full scan" if S == 1 else f"one Distribution(10000, 2, {S}, i) share",
                   "combinations": combos,
                   "algorithm": algo_name, "variant": search.last_stats.variant},
        "roofline": roofline,
        "top1": {"combination": list(top[0][0]), "mi": top[0][1]} if top else None,
    }
    if not args.no_cpu and top:
        # the reference's FP32 MI and an FP64 MI of the headline top list
        line["mi_error"] = mi_error(ds, geno, np.array([t[0] for t in top], np.int32),
                                    np.array([t[1] for t in top], np.float32))
    ds.close()
    if not args.no_order2:
        line["config1"] = bench_config1()
        line["order2"] = bench_order2()
    if not args.no_high:
        line["high_orders"] = bench_high_orders()
    if not args.no_tune and not (args.no_batch and args.no_serve):
        line["tuning"] = bench_tune()
    if not args.no_batch:
        line["batched"] = bench_batch(args.batch_datasets)
    if not args.no_serve:
        line["serving"] = bench_serve(args.serve_requests, min(args.serve_tables, args.serve_requests))
    if not (args.no_cpu and args.no_ingest):
        with tempfile.TemporaryDirectory(dir="/dev/shm" if os.path.isdir("/dev/shm") else None) as tmp:
            path = os.path.join(tmp, "bench.raw")
            write_raw(path, geno, N_CASES)
            if not args.no_ingest:
                line["ingest"] = bench_ingest(path, geno, words)
            del geno
            if not args.no_cpu:
                threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
                print(f"[bench] CPU baseline on {threads} threads ...", file=sys.stderr, flush=True)
                line["cpu_baseline"] = cpu_baseline(words, path, cw, tw, threads, args.cpu_sample_step)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

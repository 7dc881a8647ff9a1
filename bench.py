"""Benchmark: exhaustive order-3 epistasis scan on one MI355X.

Workload (the headline configuration): 10,000 SNPs x 16,000 samples (8,000
cases / 8,000 controls), HWE genotypes with MAF ~ U(0.05, 0.5), seed 1, one
planted triple (17, 4242, 9001) with a random 27-cell penetrance table in
[0.3, 0.7] from seed 2; top-100 triples by mutual information.

A step scans one slice of the full scan: the combinations whose pair prefix
has lexicographic rank = i (mod SLICES), i.e. Distribution(n, 2, SLICES, i);
--steps K scans slices 0..K-1 (K = SLICES covers every triple once).  The
metric is genotype-combination x samples per second over the timed steps.

Also reported on the same JSON line:
  roofline      the scan kernel against the int32 VALU peak (AND + popcount),
                algorithmic ops per combination x sample = 2*3^k/32 (k = 3:
                1.6875), durations from HIP events recorded on the library's
                stream around every scan launch;
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
ORDER, OUTPUTS, SLICES = 3, 100, 64
VALU_PEAK = 256 * 64 * 2.4e9          # int32 VALU lane-ops per second (MI355X)
# Bytes fetched beyond L2 per triple by the order-3 marginal scan kernel:
# rocprofv3 FETCH_SIZE per dispatch x 2.0 (the factor a plain 1 GiB copy
# shows) + WRITE_SIZE x 1.0, divided by the triples of those dispatches
# (profiles/k_scan3m_order3_10000x16000.md).
TRAFFIC_BYTES_PER_TRIPLE = {"marginal": 1563.0}


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


def cpu_baseline(words, geno, cw, tw, threads, sample_step):
    """Reference ThreadedSearch (or the oracle port) on prefix sample 0."""
    import oracle
    combos = slice_combos(N_SNPS, ORDER, sample_step, 0)
    ref = oracle.ref("avx512")
    if ref is not None and oracle.has_avx512():
        with tempfile.TemporaryDirectory(dir="/dev/shm" if os.path.isdir("/dev/shm") else None) as tmp:
            path = os.path.join(tmp, "bench.raw")
            write_raw(path, geno, N_CASES)
            d = ref.load(path)
            del geno
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
            "kind": kind,
            "sample": f"Distribution({N_SNPS}, {ORDER - 1}, {sample_step}, 0) of the same dataset "
                      f"(read from a PLINK .raw by the reference loader): "
                      f"{combos} of {math.comb(N_SNPS, ORDER)} triples, {sec:.1f} s, "
                      f"ThreadedSearch({threads}) default AVX-512 build" if kind == "reference" else
                      f"Distribution({N_SNPS}, {ORDER - 1}, {sample_step}, 0): {combos} triples, {sec:.1f} s, "
                      f"oracle AVX-512 restatement, {threads} threads"}


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
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--algorithm", default="auto", choices=["auto", "direct", "marginal"])
    ap.add_argument("--cpu-sample-step", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
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
    if args.steps < 1 or args.steps + args.warmup > SLICES:
        sys.exit(f"--steps + --warmup must be within 1..{SLICES}")

    import fiuncho_amd as fa

    algo = {"auto": fa.ALGO_AUTO, "direct": fa.ALGO_DIRECT, "marginal": fa.ALGO_MARGINAL}[args.algorithm]
    t0 = time.perf_counter()
    spec = fa.synth_spec(N_SNPS, N_CASES, N_CTRLS, seed=1, planted=PLANTED, penetrance_seed=2)
    words, geno, cw, tw = fa.synth_generate(spec, words=True, geno=not args.no_cpu)
    ds = fa.Dataset.from_words(words, N_SNPS, cw, tw, N_CASES, N_CTRLS)
    print(f"[bench] dataset generated and uploaded in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    search = fa.GpuSearch(algorithm=algo)
    # warmup on the last slices (never timed), then the timed slices 0..K-1
    for w in range(args.warmup):
        search.run_arrays(ds, ORDER, fa.Distribution(N_SNPS, 2, SLICES, SLICES - 1 - w), OUTPUTS)
    fa.lib()  # library loaded
    results = []
    combos = 0
    scan_ms = 0.0
    launches = 0
    used_algo = None
    t_start = time.perf_counter()
    for i in range(args.steps):
        c, m = search.run_arrays(ds, ORDER, fa.Distribution(N_SNPS, 2, SLICES, i), OUTPUTS)
        st = search.last_stats
        scan_ms += st.scan_ms
        launches += st.launches
        used_algo = st.algorithm
        combos += slice_combos(N_SNPS, ORDER, SLICES, i)
        results += list(zip(map(tuple, c.tolist()), m.tolist()))
        print(f"[bench] step {i} done: {st.scan_ms / 1e3:.2f} s device", file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t_start
    results.sort(key=lambda r: (-r[1], r[0]))
    top = results[:OUTPUTS]
    unit_samples = N_CASES + N_CTRLS
    value = combos * unit_samples / elapsed
    ops = combos * alg_ops_per_combo(ORDER, N_CASES, N_CTRLS)
    achieved = ops / (scan_ms / 1e3)
    algo_name = {1: "direct", 2: "marginal"}.get(used_algo, str(used_algo))
    # operations the kernel executes per triple: AND + popcount of the 8
    # (marginal) or 27 (direct) cells over every stored 32-bit word
    words32 = 2 * (cw + tw)
    executed = combos * 2 * (8 if algo_name == "marginal" else 27) * words32
    per_triple = TRAFFIC_BYTES_PER_TRIPLE.get(algo_name)
    traffic = per_triple * combos / max(launches, 1) if per_triple else None
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
                   "slices": f"{args.steps} of {SLICES} Distribution(10000, 2, {SLICES}, i) slices",
                   "combinations": combos,
                   "algorithm": algo_name},
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": VALU_PEAK / 1e12, "unit": "TOP/s",
                     "frac": achieved / VALU_PEAK, "traffic": traffic,
                     "kernel": "k_scan3m" if algo_name == "marginal" else "k_scan_direct<3>",
                     "ops_per_combo_sample": alg_ops_per_combo(ORDER, N_CASES, N_CTRLS) / unit_samples,
                     "executed_frac": executed / (scan_ms / 1e3) / VALU_PEAK,
                     "launches": launches, "avg_launch_ms": scan_ms / max(launches, 1)},
        "top1": {"combination": list(top[0][0]), "mi": top[0][1]} if top else None,
    }
    if not args.no_cpu:
        threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        print(f"[bench] CPU baseline on {threads} threads ...", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(words, geno, cw, tw, threads, args.cpu_sample_step)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

"""Write tests/golden/config_search.json: reference and oracle top lists at
the benchmark configurations' full sizes.

Runs the reference (UDC-GAC/fiuncho) compiled from its own sources by
oracle/Makefile (oracle/_ref/libfiuncho_ref.so, default AVX-512 build,
ALIGN=64) and the oracle (oracle/liboracle.so), and stores inputs' hashes and
outputs only.  Every dataset is the seeded generator's, written as a PLINK
.raw that the reference's own loader reads (Dataset<uint64_t>::read); when a
class size is a multiple of 64 that loader re-writes one case word from an
emptied buffer, so its tables differ from the generator's in that word, and
the GPU tests rebuild exactly the loader's tables
(fiu_dataset_from_genotypes with FIU_GENO_READER_FLUSH), checked against the
SHA-256 stored here.

  config1  1,000 SNPs x 4,000 samples (2,000 / 2,000), seed 1, order 2:
           ThreadedSearch(8) top-10 of the whole scan.
  config2  10,000 x 16,000 (8,000 / 8,000), seed 1, planted (17, 4242, 9001)
           with penetrance seed 2, order 3:
           ref_union   ThreadedSearch(8).run(Distribution(10000, 2, 256, 0))
                       top-100: its threads scan layer(8, t) = (2048, t),
                       t < 8, i.e. the prefixes of rank r with r mod 2048 < 8;
           oracle_union the oracle's FP64 ranking (MI rounded to float, ties
                       by combination) of the same prefixes;
           oracle_share the oracle's FP64 ranking of Distribution(10000, 2,
                       256, 0) itself (ranks r = 0 mod 256), the set one
                       Search::run call with that distribution scans
                       (ThreadedSearch(1)), as the union of the shares
                       (2048, 256 u), u < 8.
  config3  50,000 x 32,000 (16,000 / 16,000), seed 1, 20% of SNPs with MAF
           in U(0.01, 0.05), order 2:
           ref_union    ThreadedSearch(8).run(Distribution(50000, 1, 16, 0))
                        top-1,000 (prefixes r mod 128 < 8);
           oracle_union the oracle's FP64 ranking of the same prefixes.

Usage:  python tests/golden/make_config_golden.py   (needs /root/reference
built into oracle/_ref and the fiuncho_amd library for the host-only
generator; about 15 minutes on 8 cores).
"""
import hashlib
import json
import os
import sys
import tempfile
import threading
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
import fiuncho_amd as fa  # noqa: E402
from bench import write_raw  # noqa: E402  (the benchmark's PLINK .raw writer)

THREADS = 8


def sha(words):
    return hashlib.sha256(np.ascontiguousarray(words, np.uint64).tobytes()).hexdigest()


def top(combos, mi):
    return {"combos": np.asarray(combos).tolist(), "mi": [float(x) for x in mi]}


def merge(lists, n):
    """Union of per-share top lists ordered by (MI desc, combination asc)."""
    rows = []
    for c, m in lists:
        rows += [(tuple(int(x) for x in cc), float(v)) for cc, v in zip(c.tolist(), m)]
    rows.sort(key=lambda r: (-np.float32(r[1]), r[0]))
    rows = rows[:n]
    return np.array([r[0] for r in rows], np.int32), np.array([r[1] for r in rows], np.float32)


def oracle_shares(od, order, outputs, shares):
    """Oracle top lists of single-residue shares, one host thread each
    (ctypes releases the GIL during the call)."""
    out = [None] * len(shares)

    def run(i, step, offset):
        c, m, _ = oracle.search(od, order, outputs, step=step, offset=offset, threads=1, mi_mode=2)
        out[i] = (c, m)

    ths = [threading.Thread(target=run, args=(i, s, o)) for i, (s, o) in enumerate(shares)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    return out


def load_both(ref, tmp, name, geno, cases):
    path = os.path.join(tmp, name + ".raw")
    t0 = time.time()
    write_raw(path, geno, cases)
    rd = ref.load(path)
    od = oracle.read_raw(path)
    os.unlink(path)
    rw = rd.words()
    assert np.array_equal(rw, od.words), "oracle loader differs from the reference loader"
    print(f"[golden] {name}: .raw written and read by both loaders in {time.time() - t0:.1f} s", flush=True)
    return rd, od, sha(rw)


def main():
    ref = oracle.ref("avx512")
    if ref is None:
        sys.exit("oracle/_ref is not built (make -C oracle ref)")
    out = {"build": "g++ -O3 -march=x86-64-v4 -mavx512vpopcntdq -std=gnu++17 -DALIGN=64",
           "threads": THREADS}
    with tempfile.TemporaryDirectory(dir="/dev/shm" if os.path.isdir("/dev/shm") else None) as tmp:
        # config 1
        spec = fa.synth_spec(1000, 2000, 2000, seed=1)
        words, geno, _, _ = fa.synth_generate(spec, words=True, geno=True)
        rd, od, h = load_both(ref, tmp, "config1", geno, 2000)
        c, m, sec = rd.search(2, 10, threads=THREADS)
        out["config1"] = {"snps": 1000, "cases": 2000, "ctrls": 2000, "seed": 1, "order": 2, "outputs": 10,
                          "sha256": h, "same_as_generated": bool(np.array_equal(words, rd.words())),
                          "ref": top(c, m), "ref_seconds": sec}
        del rd, od, words, geno
        print("[golden] config1 done", flush=True)

        # config 2
        spec = fa.synth_spec(10000, 8000, 8000, seed=1, planted=(17, 4242, 9001), penetrance_seed=2)
        words, geno, _, _ = fa.synth_generate(spec, words=True, geno=True)
        rd, od, h = load_both(ref, tmp, "config2", geno, 8000)
        differs = int(np.count_nonzero(words != rd.words()))
        del words, geno
        c, m, sec = rd.search(3, 100, step=256, offset=0, threads=THREADS)
        print(f"[golden] config2 reference union: {sec:.1f} s", flush=True)
        t0 = time.time()
        oc, om, ev = oracle.search(od, 3, 100, step=256, offset=0, threads=THREADS, mi_mode=2)
        print(f"[golden] config2 oracle union: {time.time() - t0:.1f} s ({ev} triples)", flush=True)
        t0 = time.time()
        sc, sm = merge(oracle_shares(od, 3, 100, [(2048, 256 * u) for u in range(8)]), 100)
        print(f"[golden] config2 oracle share: {time.time() - t0:.1f} s", flush=True)
        out["config2"] = {"snps": 10000, "cases": 8000, "ctrls": 8000, "seed": 1, "planted": [17, 4242, 9001],
                          "penetrance_seed": 2, "order": 3, "outputs": 100, "sha256": h,
                          "words_differing_from_generated": differs,
                          "union": {"dist": [256, 0], "layers": [[2048, t] for t in range(8)], "triples": int(ev)},
                          "ref_union": top(c, m), "ref_seconds": sec,
                          "oracle_union": top(oc, om), "oracle_share": top(sc, sm)}
        del rd, od
        print("[golden] config2 done", flush=True)

        # config 3
        spec = fa.synth_spec(50000, 16000, 16000, seed=1, low_frac=0.2)
        words, geno, _, _ = fa.synth_generate(spec, words=True, geno=True)
        rd, od, h = load_both(ref, tmp, "config3", geno, 16000)
        differs = int(np.count_nonzero(words != rd.words()))
        del words, geno
        c, m, sec = rd.search(2, 1000, step=16, offset=0, threads=THREADS)
        print(f"[golden] config3 reference union: {sec:.1f} s", flush=True)
        t0 = time.time()
        oc, om, ev = oracle.search(od, 2, 1000, step=16, offset=0, threads=THREADS, mi_mode=2)
        print(f"[golden] config3 oracle union: {time.time() - t0:.1f} s ({ev} pairs)", flush=True)
        out["config3"] = {"snps": 50000, "cases": 16000, "ctrls": 16000, "seed": 1, "low_frac": 0.2, "order": 2,
                          "outputs": 1000, "sha256": h, "words_differing_from_generated": differs,
                          "union": {"dist": [16, 0], "layers": [[128, t] for t in range(8)], "pairs": int(ev)},
                          "ref_union": top(c, m), "ref_seconds": sec, "oracle_union": top(oc, om)}
        del rd, od
    with open(os.path.join(HERE, "config_search.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("[golden] config_search.json written", flush=True)


if __name__ == "__main__":
    main()

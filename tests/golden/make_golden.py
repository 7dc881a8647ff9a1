"""Write the golden vectors of tests/golden/ from the reference itself.

Runs the reference (UDC-GAC/fiuncho) compiled from its own sources by
oracle/Makefile (oracle/_ref/libfiuncho_ref.so: default AVX-512 build,
libfiuncho_ref_base.so: scalar build, both ALIGN=64) and stores inputs and
outputs only:

  testdata_tables.npz   every combination of tests/data (10 SNPs) at orders
                        2 to 8: counts (cases, controls; 3^k cells) and the
                        float MI of both reference builds
  testdata_search.json  ThreadedSearch(1).run top-90 at orders 2 to 9, and the
                        SHA-256 of the bit tables the reference loader builds
                        from the .tped/.tfam, .raw and GAMETES .raw files
  loader_quirk.npz      a small .raw and .tped/.tfam (their text is stored) whose
                        sample order triggers the fill loop's re-flush of a full
                        word, and the bit tables the reference loader builds
  synth_search.json     ThreadedSearch(8).run on small synthetic datasets made
                        by fiuncho_amd's seeded generator (written as PLINK
                        .raw files so that the reference's own loader reads
                        them), with the SHA-256 of the reference's bit tables

Usage:  python tests/golden/make_golden.py   (needs /root/reference and a
built fiuncho_amd library for the host-only generator).
"""
import hashlib
import itertools
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
import fiuncho_amd as fa  # noqa: E402

DATA = os.path.join(HERE, "data")

# small synthetic cases: (name, snps, cases, ctrls, seed, order, outputs, planted)
SYNTH = [
    ("s2_ragged", 64, 101, 77, 11, 2, 20, ()),
    ("s2_small", 200, 500, 500, 12, 2, 50, (3, 150)),
    ("s3_small", 60, 300, 260, 13, 3, 40, (5, 17, 40)),
    ("s3_ragged", 45, 129, 63, 14, 3, 100, ()),
    ("s4_small", 24, 200, 200, 15, 4, 30, (1, 7, 12, 20)),
    ("s5_small", 16, 300, 260, 16, 5, 40, (2, 5, 9, 11, 14)),
    ("s6_small", 12, 400, 350, 17, 6, 60, (0, 3, 4, 7, 9, 11)),
    ("s7_small", 11, 300, 280, 18, 7, 50, ()),
    ("s8_small", 10, 260, 250, 19, 8, 40, ()),
    ("s9_small", 11, 240, 230, 20, 9, 30, ()),
]


def sha(words):
    return hashlib.sha256(np.ascontiguousarray(words, np.uint64).tobytes()).hexdigest()


def write_raw(path, geno, cases):
    """PLINK .raw with FID IID PAT MAT SEX PHENOTYPE then one column per SNP."""
    snps, n = geno.shape
    with open(path, "w") as f:
        f.write("FID IID PAT MAT SEX PHENOTYPE " + " ".join(f"S{i}_A" for i in range(snps)) + "\n")
        for j in range(n):
            ph = 2 if j < cases else 1
            f.write(f"i{j} i{j} 0 0 0 {ph} " + " ".join(str(int(v)) for v in geno[:, j]) + "\n")


def main():
    ref = oracle.ref("avx512")
    base = oracle.ref("base")
    if ref is None or base is None:
        sys.exit("oracle/_ref is not built (make -C oracle ref)")
    tp = ref.load(os.path.join(DATA, "test.tped"), os.path.join(DATA, "test.tfam"))
    tb = base.load(os.path.join(DATA, "test.tped"), os.path.join(DATA, "test.tfam"))
    raw = ref.load(os.path.join(DATA, "test.raw"))
    gam = ref.load(os.path.join(DATA, "test.gametes.raw"))
    tables = {}
    for k in (2, 3, 4, 5, 6, 7, 8):
        combos = np.array(list(itertools.combinations(range(tp.snps), k)), np.int32)
        cc = np.zeros((len(combos), 3 ** k), np.uint32)
        ct = np.zeros_like(cc)
        mi512 = np.zeros(len(combos), np.float32)
        mib = np.zeros(len(combos), np.float32)
        for i, c in enumerate(combos):
            cc[i], ct[i], mi512[i] = tp.combination(c)
            _, _, mib[i] = tb.combination(c)
        tables[f"k{k}_combos"] = combos
        tables[f"k{k}_cases"] = cc
        tables[f"k{k}_ctrls"] = ct
        tables[f"k{k}_mi_avx512"] = mi512
        tables[f"k{k}_mi_base"] = mib
    pair = tp.combine_pair(0, 1)
    tables["combine_0_1"] = pair
    np.savez_compressed(os.path.join(HERE, "testdata_tables.npz"), **tables)

    search = {"shape": {"snps": tp.snps, "cases": tp.cases, "ctrls": tp.ctrls,
                        "cases_words": tp.cases_words, "ctrls_words": tp.ctrls_words},
              "sha256": {"tped": sha(tp.words()), "raw": sha(raw.words()), "gametes": sha(gam.words())},
              "build": "g++ -O3 -march=x86-64-v4 -mavx512vpopcntdq -std=gnu++17 -DALIGN=64",
              "top": {}}
    for k in (2, 3, 4, 5, 6, 7, 8, 9):
        combos, mi, _ = tp.search(k, 90, threads=1)
        search["top"][str(k)] = {"combos": combos.tolist(), "mi": [float(x) for x in mi]}
    with open(os.path.join(HERE, "testdata_search.json"), "w") as f:
        json.dump(search, f, indent=1)

    # loader fixture whose sample order makes the reference's fill loop
    # re-flush a full word (64 cases then a control, 64 controls then a case)
    rng = np.random.default_rng(77)
    order = [0] * 64 + [1] + [0] * 67 + [1] * 63 + [0] + [1] * 70 + [0] * 60
    geno = rng.integers(0, 3, size=(5, len(order))).astype(np.int8)
    quirk = {}
    with tempfile.TemporaryDirectory() as tmp:
        raw = os.path.join(tmp, "q.raw")
        with open(raw, "w") as f:
            f.write("FID IID PAT MAT SEX PHENOTYPE " + " ".join(f"Q{i}_T" for i in range(5)) + "\n")
            for j, cl in enumerate(order):
                f.write(f"f{j} i{j} 0 0 1 {1 if cl else 2} " + " ".join(str(int(v)) for v in geno[:, j]) + "\n")
        tped, tfam = os.path.join(tmp, "q.tped"), os.path.join(tmp, "q.tfam")
        with open(tfam, "w") as f:
            for j, cl in enumerate(order):
                f.write(f"f{j} i{j} 0 0 1 {1 if cl else 2}\n")
        pairs = {0: "G G", 1: "A G", 2: "A A"}
        with open(tped, "w") as f:
            for s in range(5):
                f.write(f"1 q{s} 0 {100 + s} " + " ".join(pairs[int(v)] for v in geno[s]) + "\n")
        for name, files in (("raw", (raw,)), ("tped", (tped, tfam))):
            d = ref.load(*files)
            quirk[f"{name}_words"] = d.words()
            quirk[f"{name}_shape"] = np.array([d.snps, d.cases, d.ctrls, d.cases_words, d.ctrls_words])
            with open(files[0]) as fh:
                quirk[f"{name}_text"] = np.frombuffer(fh.read().encode(), np.uint8)
        with open(tfam) as fh:
            quirk["tfam_text"] = np.frombuffer(fh.read().encode(), np.uint8)
    np.savez_compressed(os.path.join(HERE, "loader_quirk.npz"), **quirk)

    synth = {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, snps, cases, ctrls, seed, order, outputs, planted in SYNTH:
            spec = fa.synth_spec(snps, cases, ctrls, seed=seed, planted=planted)
            words, geno, cw, tw = fa.synth_generate(spec, words=True, geno=True)
            path = os.path.join(tmp, name + ".raw")
            write_raw(path, geno, cases)
            d = ref.load(path)
            assert d.snps == snps and d.cases == cases and d.ctrls == ctrls
            combos, mi, _ = d.search(order, outputs, threads=8)
            synth[name] = {"snps": snps, "cases": cases, "ctrls": ctrls, "seed": seed,
                           "planted": list(planted), "order": order, "outputs": outputs,
                           "sha256": sha(d.words()), "combos": combos.tolist(),
                           "mi": [float(x) for x in mi]}
    with open(os.path.join(HERE, "synth_search.json"), "w") as f:
        json.dump(synth, f, indent=1)
    print("golden vectors written")


if __name__ == "__main__":
    main()

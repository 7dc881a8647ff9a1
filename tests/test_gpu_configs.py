"""GPU parity at the benchmark configurations' full sizes.

Each configuration's data is the seeded generator's, exactly as bench.py
generates it.  Reference lists come from tests/golden/config_search.json
(written by tests/golden/make_config_golden.py from the reference built from
its own sources, reading the data through its own loader, whose tables the
tests rebuild with FIU_GENO_READER_FLUSH and check by SHA-256).

Bars:
* counts bit-exact: the GPU's contingency tables of the returned
  combinations (fiu_contingency_batch) equal the oracle's prefix-chain counts
  (the reference's combine / combine_and_popcnt, ThreadedSearch.h:160-176);
* MI: every returned float is within half an ulp of the FP64 MI of those
  counts (f64_tol), and within 1e-5 |mi_ref| + 8 eps32 (H(X) + H(Y) + H(X,Y))
  of the reference's FP32 MutualInformation<float>::compute of the same
  counts (ref_tol, computed by the reference build when present, else by the
  oracle's bit-exact restatement of it);
* lists: identical to the oracle's FP64 ranking of the same combinations
  (combinations, order and float MI bit for bit); equal to the reference's
  list up to ties: a combination in only one of the two lists has an MI
  within 4e-6 of the N-th value (above every per-entry ref_tol here), and
  common combinations agree within ref_tol.
"""
import hashlib
import json
import math
import os

import numpy as np
import pytest

import fiuncho_amd as fa
import oracle as O

from .conftest import GOLDEN
from .test_gpu_parity import f64_tol, ref_tol

pytestmark = pytest.mark.gpu

_GOLD_PATH = os.path.join(GOLDEN, "config_search.json")
if not os.path.exists(_GOLD_PATH):
    pytest.skip("tests/golden/config_search.json not written (tests/golden/make_config_golden.py)",
                allow_module_level=True)
with open(_GOLD_PATH) as _f:
    GOLD = json.load(_f)

TIE_BAND = 4e-6


def sha(words):
    return hashlib.sha256(np.ascontiguousarray(words, np.uint64).tobytes()).hexdigest()


def gen(snps, cases, ctrls, geno=False, **kw):
    spec = fa.synth_spec(snps, cases, ctrls, seed=1, **kw)
    return fa.synth_generate(spec, words=True, geno=geno)


def reader_tables(geno, cases, g):
    """The tables the reference's loader builds from these genotypes."""
    is_ctrl = (np.arange(geno.shape[1]) >= cases).astype(np.uint8)
    d = fa.Dataset.from_genotypes(geno, is_ctrl, reader_flush=True)
    w = d.words()
    assert sha(w) == g["sha256"]
    return d, w


def ref_mi(cc, ct, k, cases, ctrls):
    r = O.ref("avx512")
    if r is not None and O.has_avx512():
        return r.mi(cc, ct, k, cases, ctrls)
    return O.mi_f32_ref512(cc, ct, k, cases, ctrls) if O.has_avx512() else O.mi_f32_base(cc, ct, k, cases, ctrls)


def check_counts_and_mi(d, od, combos, mi, order):
    """Counts of every returned combination against the oracle, its MI
    against FP64 and the reference's FP32; returns the error statistics."""
    counts, _ = fa.contingency_batch(d, combos)
    cells = 3 ** order
    rel, ab = [], []
    for i, c in enumerate(combos):
        cc, ct = O.counts(od, c)
        assert np.array_equal(counts[i, 0, :cells], cc) and np.array_equal(counts[i, 1, :cells], ct), c
        m64 = O.mi_f64(cc, ct, order, od.cases, od.ctrls)
        assert abs(float(mi[i]) - m64) <= f64_tol(m64), c
        r = ref_mi(cc, ct, order, od.cases, od.ctrls)
        assert abs(float(mi[i]) - r) <= ref_tol(r, cc, ct), c
        rel.append(abs(float(mi[i]) - r) / abs(r) if r else 0.0)
        ab.append(abs(float(mi[i]) - m64))
    return max(rel), float(np.median(rel)), max(ab)


def keys_sorted(combos, mi):
    keys = [(-float(v), tuple(c)) for c, v in zip(np.asarray(combos).tolist(), mi)]
    return keys == sorted(keys) and len(set(k[1] for k in keys)) == len(keys)


def identical(c, m, g):
    return np.asarray(c).tolist() == g["combos"] and np.array_equal(np.asarray(m, np.float32),
                                                                    np.float32(g["mi"]))


def same_up_to_ties(combos, mi, g):
    a = {tuple(c): float(v) for c, v in zip(np.asarray(combos).tolist(), mi)}
    b = {tuple(c): float(v) for c, v in zip(g["combos"], g["mi"])}
    if len(a) != len(b):
        return False
    last = min(b.values())
    for c in set(a) ^ set(b):
        if abs(a.get(c, b.get(c)) - last) > TIE_BAND:
            return False
    return all(abs(a[c] - b[c]) <= TIE_BAND for c in set(a) & set(b))


def merge(lists, n):
    rows = []
    for c, m in lists:
        rows += [(tuple(cc), float(v)) for cc, v in zip(c.tolist(), m)]
    rows.sort(key=lambda r: (-r[1], r[0]))
    rows = rows[:n]
    return np.array([r[0] for r in rows], np.int32), np.array([r[1] for r in rows], np.float32)


# ---------------------------------------------------------------------------
# Config 1: order 2, 1,000 SNPs x 4,000 samples, top-10
# ---------------------------------------------------------------------------
def test_config1_full_scan():
    g = GOLD["config1"]
    print("case: config 1, 1,000 x 4,000 order 2", flush=True)
    words, _, cw, tw = gen(1000, 2000, 2000)
    assert sha(words) == g["sha256"] and g["same_as_generated"]
    d = fa.Dataset.from_words(words, 1000, cw, tw, 2000, 2000)
    c, m = fa.GpuSearch().run_arrays(d, 2, fa.Distribution(1000, 1), 10)
    od = O.from_words(words, 1000, 2000, 2000, cw, tw)
    oc, om, _ = O.search(od, 2, 10, threads=8, mi_mode=2)
    assert c.tolist() == oc.tolist() and np.array_equal(m, om)
    assert same_up_to_ties(c, m, g["ref"])
    check_counts_and_mi(d, od, c, m, 2)
    d.close()


# ---------------------------------------------------------------------------
# Config 2: order 3, 10,000 SNPs x 16,000 samples, planted (17, 4242, 9001)
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def cfg2():
    print("case: config 2 generation", flush=True)
    return gen(10000, 8000, 8000, geno=True, planted=(17, 4242, 9001), penetrance_seed=2)


def prefix_rank2(a, b, n):
    return a * (2 * n - a - 1) // 2 + (b - a - 1)


def test_config2_full_scan(cfg2):
    # the benchmark's dataset and call: the whole scan on the matrix cores
    words, _, cw, tw = cfg2
    print("case: config 2 full scan (matrix cores)", flush=True)
    d = fa.Dataset.from_words(words, 10000, cw, tw, 8000, 8000)
    s = fa.GpuSearch()
    c, m = s.run_arrays(d, 3, fa.Distribution(10000, 2), 100)
    assert s.last_stats.algorithm == fa.ALGO_MFMA
    assert s.last_stats.combinations == math.comb(10000, 3)
    assert c.shape == (100, 3) and c[0].tolist() == [17, 4242, 9001]
    assert keys_sorted(c, m) and (np.diff(c, axis=1) > 0).all()
    od = O.from_words(words, 10000, 8000, 8000, cw, tw)
    rel_max, rel_med, abs_max = check_counts_and_mi(d, od, c, m, 3)
    print(f"case: config 2 MI error vs reference FP32: max {rel_max:.2e} rel, median {rel_med:.2e}; "
          f"vs FP64: max {abs_max:.2e} abs", flush=True)
    d.close()


def test_config2_reference_tables(cfg2):
    g = GOLD["config2"]
    _, geno, _, _ = cfg2
    print("case: config 2 on the reference loader's tables", flush=True)
    d, w = reader_tables(geno, 8000, g)
    od = O.from_words(w, 10000, 8000, 8000, d.cases_words, d.ctrls_words)
    # ThreadedSearch(8) on Distribution(10000, 2, 256, 0): the union of the
    # layers (2048, t), t < 8, one search each, merged
    lists = [fa.GpuSearch().run_arrays(d, 3, fa.Distribution(10000, 2, 2048, t), 100) for t in range(8)]
    uc, um = merge(lists, 100)
    assert identical(uc, um, g["oracle_union"])
    assert same_up_to_ties(uc, um, g["ref_union"])
    check_counts_and_mi(d, od, uc, um, 3)
    # one Search::run call on Distribution(10000, 2, 256, 0), on the matrix
    # cores (the share masked in the epilogue)
    print("case: config 2 share (256, 0) on the matrix cores", flush=True)
    s = fa.GpuSearch(algorithm=fa.ALGO_MFMA)
    c, m = s.run_arrays(d, 3, fa.Distribution(10000, 2, 256, 0), 100)
    assert s.last_stats.algorithm == fa.ALGO_MFMA
    assert identical(c, m, g["oracle_share"])
    # the whole scan: its entries inside the union are exactly the union
    # list's entries ranked at or above its 100th
    print("case: config 2 full scan on the reference loader's tables", flush=True)
    fc, fm = fa.GpuSearch().run_arrays(d, 3, fa.Distribution(10000, 2), 100)
    assert fc[0].tolist() == [17, 4242, 9001]
    last = (-float(fm[-1]), tuple(fc[-1].tolist()))
    inside = [(tuple(x), float(v)) for x, v in zip(fc.tolist(), fm) if prefix_rank2(x[0], x[1], 10000) % 2048 < 8]
    want = [(tuple(x), v) for x, v in zip(g["oracle_union"]["combos"], np.float32(g["oracle_union"]["mi"]).tolist())
            if (-v, tuple(x)) <= last]
    assert inside == want
    d.close()


# ---------------------------------------------------------------------------
# Config 3: order 2, 50,000 SNPs x 32,000 samples, top-1,000
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def cfg3():
    print("case: config 3 generation", flush=True)
    return gen(50000, 16000, 16000, geno=True, low_frac=0.2)


def test_config3_operand_formats(cfg3):
    words, _, cw, tw = cfg3
    d = fa.Dataset.from_words(words, 50000, cw, tw, 16000, 16000)
    dist = fa.Distribution(50000, 1)
    res = {}
    for name, algo, int8 in (("fp4", fa.ALGO_MFMA, False), ("int8", fa.ALGO_MFMA, True),
                             ("popcount", fa.ALGO_MARGINAL, False)):
        print(f"case: config 3 full scan, {name}", flush=True)
        s = fa.GpuSearch(algorithm=algo, mfma_int8=int8)
        res[name] = s.run_arrays(d, 2, dist, 1000)
        assert s.last_stats.algorithm == algo
    c, m = res["fp4"]
    for k in ("int8", "popcount"):
        assert c.tolist() == res[k][0].tolist() and np.array_equal(m, res[k][1]), k
    assert c.shape == (1000, 2) and keys_sorted(c, m)
    od = O.from_words(words, 50000, 16000, 16000, cw, tw)
    rel_max, rel_med, abs_max = check_counts_and_mi(d, od, c, m, 2)
    print(f"case: config 3 MI error vs reference FP32: max {rel_max:.2e} rel, median {rel_med:.2e}; "
          f"vs FP64: max {abs_max:.2e} abs", flush=True)
    d.close()


def test_config3_reference_tables(cfg3):
    g = GOLD["config3"]
    _, geno, _, _ = cfg3
    print("case: config 3 on the reference loader's tables", flush=True)
    d, w = reader_tables(geno, 16000, g)
    # ThreadedSearch(8) on Distribution(50000, 1, 16, 0): layers (128, t), t < 8
    lists = [fa.GpuSearch().run_arrays(d, 2, fa.Distribution(50000, 1, 128, t), 1000) for t in range(8)]
    uc, um = merge(lists, 1000)
    assert identical(uc, um, g["oracle_union"])
    assert same_up_to_ties(uc, um, g["ref_union"])
    od = O.from_words(w, 50000, 16000, 16000, d.cases_words, d.ctrls_words)
    check_counts_and_mi(d, od, uc[:200], um[:200], 2)
    d.close()


# ---------------------------------------------------------------------------
# Config 4: 64 uneven datasets in one batched call (the benchmark's scaled form)
# ---------------------------------------------------------------------------
def test_config4_batch():
    import bench
    specs = bench.uneven_specs(64, 100, bench.BATCH_SCALE)
    print("case: config 4, 64 uneven datasets", flush=True)
    dss, small = [], []
    for i, sp in enumerate(specs):
        words, cw, tw = bench.make_words(sp)
        dss.append(fa.Dataset.from_words(words, sp["snps"], cw, tw, sp["cases"], sp["ctrls"]))
        # the smallest searches of each order are also checked against the oracle
        small.append((bench.combo_samples(sp), i, words if bench.combo_samples(sp) < 1e11 else None, cw, tw))
        del words
    scans = [(d, sp["order"], fa.Distribution(sp["snps"], sp["order"] - 1), bench.BATCH_OUTPUTS)
             for d, sp in zip(dss, specs)]
    plain = [fa.GpuSearch().run_arrays(*s) for s in scans]
    batched, st = fa.search_batch(scans)
    assert st.steps > 0
    for i, (a, b) in enumerate(zip(plain, batched)):
        assert a[0].tolist() == b[0].tolist() and np.array_equal(a[1], b[1]), i
        assert len(a[1]) == bench.BATCH_OUTPUTS
    picked = {}
    for work, i, words, cw, tw in sorted(small, key=lambda t: t[0]):
        o = specs[i]["order"]
        if words is not None and picked.get(o, 0) < 2:
            picked[o] = picked.get(o, 0) + 1
            sp = specs[i]
            print(f"case: config 4 dataset {i} (order {o}, {sp['snps']} SNPs) against the oracle", flush=True)
            od = O.from_words(words, sp["snps"], sp["cases"], sp["ctrls"], cw, tw)
            oc, om, _ = O.search(od, o, bench.BATCH_OUTPUTS, threads=8, mi_mode=2)
            assert batched[i][0].tolist() == oc.tolist() and np.array_equal(batched[i][1], om), i
    assert picked.get(2, 0) == 2 and picked.get(3, 0) == 2
    for d in dss:
        d.close()

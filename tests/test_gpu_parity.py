"""GPU parity: the HIP path through the C ABI against the oracle, the
reference's own fixtures and the golden vectors written by the reference.

Bars: counts bit-exact; MI within
    |mi_gpu - mi_ref| <= 1e-5 |mi_ref| + 8 eps32 (H(X) + H(Y) + H(X,Y))
of the reference's FP32 value, and within one float rounding
(0.5 ulp + 1e-12) of the FP64 value of the exact counts; top-N sets equal up
to entries tied with the N-th within that tolerance.
"""
import itertools
import json
import math
import os

import numpy as np
import pytest

import fiuncho_amd as fa
import oracle as O

from .conftest import DATA, GOLDEN
from .test_oracle import matches_mpi3snp

pytestmark = pytest.mark.gpu

EPS32 = 2.0 ** -23
TPED, TFAM = os.path.join(DATA, "test.tped"), os.path.join(DATA, "test.tfam")


def entropies(cc, ct):
    cc, ct = np.asarray(cc, float), np.asarray(ct, float)
    n = cc.sum() + ct.sum()

    def h(v):
        v = v[v > 0] / n
        return float(-(v * np.log(v)).sum())

    return h(cc + ct) + h(np.array([cc.sum(), ct.sum()])) + h(np.r_[cc, ct])


def ref_tol(mi_ref, cc, ct):
    return 1e-5 * abs(mi_ref) + 8 * EPS32 * entropies(cc, ct)


def f64_tol(m64):
    return 0.5 * float(np.spacing(np.float32(abs(m64)))) + 1e-12


@pytest.fixture(scope="module")
def tables():
    return np.load(os.path.join(GOLDEN, "testdata_tables.npz"))


@pytest.fixture(scope="module")
def search_golden():
    with open(os.path.join(GOLDEN, "testdata_search.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def ds():
    print("case: tests/data tped dataset upload", flush=True)
    return fa.Dataset.read([TPED, TFAM])


def test_dataset_read_all_formats(ds, search_golden):
    import hashlib
    assert (ds.snps, ds.cases, ds.ctrls, ds.cases_words, ds.ctrls_words) == (10, 9000, 9800, 144, 160)
    for files in ([TFAM, TPED], [os.path.join(DATA, "test.raw")], [os.path.join(DATA, "test.gametes.raw")]):
        d = fa.Dataset.read(files)
        assert hashlib.sha256(d.words().tobytes()).hexdigest() == search_golden["sha256"]["tped"]


def test_dataset_read_errors():
    with pytest.raises(fa.FiuError):
        fa.Dataset.read([os.path.join(DATA, "test.tped")])
    with pytest.raises(fa.FiuError):
        fa.Dataset.read([TPED, TPED])
    with pytest.raises(fa.FiuError):
        fa.Dataset.read(["/nonexistent.raw"])


def test_contingency_all_combinations(ds, tables):
    print("case: contingency_batch all 1,002 combinations", flush=True)
    for k in (2, 3, 4, 5, 6, 7, 8):
        combos = tables[f"k{k}_combos"]
        counts, mi = fa.contingency_batch(ds, combos)
        cells = 3 ** k
        assert np.array_equal(counts[:, 0, :cells], tables[f"k{k}_cases"])
        assert np.array_equal(counts[:, 1, :cells], tables[f"k{k}_ctrls"])
        assert (counts[:, :, cells:] == 0).all()
        for i in range(len(combos)):
            cc, ct = tables[f"k{k}_cases"][i], tables[f"k{k}_ctrls"][i]
            m64 = O.mi_f64(cc, ct, k, 9000, 9800)
            assert abs(mi[i] - m64) <= f64_tol(m64)
            r = tables[f"k{k}_mi_avx512"][i]
            assert abs(float(mi[i]) - float(r)) <= ref_tol(r, cc, ct)


def test_combine_batch(ds, tables):
    out = fa.combine_batch(ds, np.array([[0, 1]]))
    assert np.array_equal(out[0], tables["combine_0_1"])
    od = O.read_tped(TPED, TFAM)
    cw, tw = od.cases_words, od.ctrls_words
    for combo in ([3], [0, 5, 9], [1, 2, 6, 8]):
        got = fa.combine_batch(ds, np.array([combo]))[0]
        pc, pt = od.snp(combo[0])
        size = 3
        for s in combo[1:]:
            c2, t2 = od.snp(s)
            nc, nt = np.zeros(size * 3 * cw, np.uint64), np.zeros(size * 3 * tw, np.uint64)
            O.lib().orc_combine(np.ascontiguousarray(pc), size, np.ascontiguousarray(c2), cw, nc)
            O.lib().orc_combine(np.ascontiguousarray(pt), size, np.ascontiguousarray(t2), tw, nt)
            pc, pt, size = nc, nt, size * 3
        assert np.array_equal(got, np.r_[pc, pt]), combo


def test_mi_batch_known_answers():
    # tests/mi.cpp:14-30
    t = np.zeros((2, 2, 16), np.uint32)
    t[0, 0, :9], t[0, 1, :9] = 1, 2
    t[1, 0, 0], t[1, 1, 1] = 9, 18
    mi = fa.mi_batch(t, 2, 9, 18)
    assert abs(mi[0]) < 1e-5
    assert abs(mi[1] - 0.6365141682948129) < 1e-5


@pytest.mark.parametrize("algorithm", [fa.ALGO_AUTO, fa.ALGO_DIRECT])
def test_search_testdata(ds, search_golden, tables, algorithm):
    # tests/threadedsearch.cpp:17-31 plus the reference's own top-90
    print(f"case: search tests/data algorithm={algorithm}", flush=True)
    for o in ((2, 3, 4, 5, 6, 7, 8) if algorithm == fa.ALGO_AUTO else (2, 3, 4, 7, 8)):   # 5-6: marginal only
        res = fa.GpuSearch(algorithm=algorithm).run(ds, o, fa.Distribution(10, o - 1), 90)
        combos = np.array([r.combination for r in res])
        mi = np.array([r.val for r in res], np.float32)
        assert len(res) > 0
        if o > 2:
            assert len(res) == min(90, math.comb(10, o))
        assert combos.shape[1] == o
        assert len({tuple(c) for c in combos}) == len(combos)
        assert (np.diff(combos, axis=1) > 0).all()
        if o == 3:
            assert matches_mpi3snp(combos, mi)
        g = search_golden["top"][str(o)]
        assert {tuple(c) for c in combos} == {tuple(c) for c in g["combos"]}
        # order: MI descending, ties by combination
        keys = [(-float(v), tuple(c)) for c, v in zip(combos, mi)]
        assert keys == sorted(keys)
        # values against the reference and against FP64 of the exact counts
        ref = {tuple(c): v for c, v in zip(g["combos"], g["mi"])}
        idx = {tuple(c): i for i, c in enumerate(tables[f"k{o}_combos"].tolist())}
        for c, v in zip(combos, mi):
            i = idx[tuple(c)]
            cc, ct = tables[f"k{o}_cases"][i], tables[f"k{o}_ctrls"][i]
            assert abs(float(v) - ref[tuple(c)]) <= ref_tol(ref[tuple(c)], cc, ct)
            m64 = O.mi_f64(cc, ct, o, 9000, 9800)
            assert abs(float(v) - m64) <= f64_tol(m64)


def synth(snps, cases, ctrls, seed, planted=()):
    spec = fa.synth_spec(snps, cases, ctrls, seed=seed, planted=planted)
    words, geno, cw, tw = fa.synth_generate(spec, words=True, geno=True)
    d = fa.Dataset.from_words(words, snps, cw, tw, cases, ctrls)
    od = O.pack(geno, np.r_[np.zeros(cases, np.uint8), np.ones(ctrls, np.uint8)])
    return d, od


def same_topn(combos, mi, ocombos, omi, tol=1e-6):
    """Sets equal up to entries tied (within tol) with the N-th value."""
    a = {tuple(c): float(v) for c, v in zip(combos, mi)}
    b = {tuple(c): float(v) for c, v in zip(ocombos, omi)}
    if len(a) != len(b):
        return False
    if not len(b):
        return True
    last = min(b.values())
    for c in set(a) ^ set(b):
        v = a.get(c, b.get(c))
        if abs(v - last) > tol:
            return False
    for c in set(a) & set(b):
        if abs(a[c] - b[c]) > max(tol * abs(b[c]), 1e-7):
            return False
    return True


def test_search_synthetic_golden():
    with open(os.path.join(GOLDEN, "synth_search.json")) as f:
        synth_g = json.load(f)
    for name, g in synth_g.items():
        print(f"case: golden synthetic {name}", flush=True)
        d, od = synth(g["snps"], g["cases"], g["ctrls"], g["seed"], g["planted"])
        combos, mi = fa.GpuSearch().run_arrays(d, g["order"], fa.Distribution(g["snps"], g["order"] - 1), g["outputs"])
        assert same_topn(combos, mi, g["combos"], g["mi"], tol=2e-5), name
        for c, v in zip(combos, mi):
            cc, ct = O.counts(od, c)
            m64 = O.mi_f64(cc, ct, g["order"], g["cases"], g["ctrls"])
            assert abs(float(v) - m64) <= f64_tol(m64), (name, c)


@pytest.mark.parametrize("snps,cases,ctrls,order,outputs,seed", [
    (70, 129, 63, 2, 40, 21),       # ragged classes, partial words
    (128, 512, 512, 2, 100, 22),
    (50, 333, 190, 3, 60, 23),
    (23, 64, 64, 4, 50, 24),
    (5, 40, 40, 3, 100, 25),        # fewer combinations than outputs
])
def test_search_vs_oracle_fp64(snps, cases, ctrls, order, outputs, seed):
    # identical top-N (order, combinations and float values) to the oracle
    # ranking by FP64 MI rounded to float, ties by combination
    print(f"case: synthetic {snps}x{cases}+{ctrls} k={order}", flush=True)
    d, od = synth(snps, cases, ctrls, seed)
    oc, om, _ = O.search(od, order, outputs, threads=4, mi_mode=2)
    algos = (fa.ALGO_AUTO, fa.ALGO_DIRECT, fa.ALGO_MARGINAL) + ((fa.ALGO_MFMA,) if order <= 3 else ())
    for algorithm in algos:
        s = fa.GpuSearch(algorithm=algorithm)
        combos, mi = s.run_arrays(d, order, fa.Distribution(snps, order - 1), outputs)
        used = s.last_stats.algorithm
        if algorithm == fa.ALGO_DIRECT:
            assert used == fa.ALGO_DIRECT
        elif order > 3:
            assert used == fa.ALGO_MARGINAL   # orders 4-5: marginal in the last SNP
        elif algorithm != fa.ALGO_MARGINAL:
            assert used == fa.ALGO_MFMA   # the default of whole order-2 and order-3 scans
        else:
            assert used == fa.ALGO_MARGINAL
        assert combos.tolist() == oc.tolist(), (algorithm, used)
        assert np.array_equal(mi, om), (algorithm, used)
        assert s.last_stats.combinations == math.comb(snps, order)


def test_distribution_partition():
    # every combination exactly once over the offsets of a step (MPIEngine
    # rank split), each with the same MI as in the full scan
    print("case: distribution partition", flush=True)
    d, _ = synth(30, 200, 180, 31)
    total = math.comb(30, 3)
    full_c, full_m = fa.GpuSearch().run_arrays(d, 3, fa.Distribution(30, 2), 4096)
    assert len(full_c) == total
    ref = {tuple(c): v for c, v in zip(full_c.tolist(), full_m)}
    for step in (3, 7):
        seen = {}
        for off in range(step):
            dist = fa.Distribution(30, 2).layer(step, off)
            c, m = fa.GpuSearch().run_arrays(d, 3, dist, 4096)
            prefixes = O.dist_prefixes(30, 2, step, off)
            allowed = {tuple(p) for p in prefixes.tolist()}
            for cc, mm in zip(c.tolist(), m):
                assert tuple(cc[:2]) in allowed
                assert tuple(cc) not in seen
                seen[tuple(cc)] = mm
        assert seen == ref


def test_edge_cases():
    d, od = synth(12, 70, 50, 41)
    s = fa.GpuSearch()
    assert s.run(d, 3, fa.Distribution(12, 2), 0) == []
    assert len(s.run(d, 2, fa.Distribution(12, 1), 1000)) == math.comb(12, 2)
    # order above the SNP count: no combinations
    small, _ = synth(3, 30, 30, 42)
    assert s.run(small, 4, fa.Distribution(3, 3), 10) == []
    # offset past the last prefix
    assert s.run(d, 2, fa.Distribution(12, 1, 5, 40), 10) == []
    with pytest.raises(fa.FiuError):
        s.run(d, 1, fa.Distribution(12, 0), 10)
    with pytest.raises(fa.FiuError):
        s.run(d, 2, fa.Distribution(12, 1, 0, 0), 10)


def test_ties_broken_by_combination():
    # identical SNPs give identical MI: order must be lexicographic
    spec = fa.synth_spec(1, 100, 90, seed=51)
    w1, g1, cw, tw = fa.synth_generate(spec, words=True, geno=True)
    words = np.tile(w1, 8)
    d = fa.Dataset.from_words(words, 8, cw, tw, 100, 90)
    res = fa.GpuSearch().run(d, 2, fa.Distribution(8, 1), 28)
    assert [r.combination for r in res] == [list(c) for c in itertools.combinations(range(8), 2)]
    assert len({r.val for r in res}) == 1


def test_incomplete_genotypes_use_exact_counts():
    # genotype values outside {0,1,2} leave a sample out of every row
    rng = np.random.default_rng(61)
    geno = rng.integers(0, 3, size=(20, 150)).astype(np.int8)
    geno[rng.random((20, 150)) < 0.05] = 3
    is_ctrl = (np.arange(150) >= 80).astype(np.uint8)
    od = O.pack(geno, is_ctrl)
    d = fa.Dataset.from_words(od.words, 20, od.cases_words, od.ctrls_words, 80, 70)
    for order in (2, 3):
        combos, mi = fa.GpuSearch().run_arrays(d, order, fa.Distribution(20, order - 1), 50)
        oc, om, _ = O.search(od, order, 50, threads=2, mi_mode=2)
        assert combos.tolist() == oc.tolist() and np.array_equal(mi, om)
    with pytest.raises(fa.FiuError):
        fa.GpuSearch(algorithm=fa.ALGO_MARGINAL).run(d, 2, fa.Distribution(20, 1), 5)


def test_uncalled_sample_keeps_exact_mi():
    # one sample per class uncalled at every SNP: every SNP covers the same
    # samples (the marginal and matrix-core scans apply) but the tables sum
    # to less than the class sizes, so the MI is over sub-probabilities and
    # the chi-square upper bound must not be used to reject combinations
    rng = np.random.default_rng(62)
    geno = rng.integers(0, 3, size=(40, 300)).astype(np.int8)
    geno[:, 7] = 3
    geno[:, 250] = 3
    is_ctrl = (np.arange(300) >= 160).astype(np.uint8)
    od = O.pack(geno, is_ctrl)
    d = fa.Dataset.from_words(od.words, 40, od.cases_words, od.ctrls_words, 160, 140)
    for order in (2, 3, 4, 5, 6):
        print(f"case: uncalled sample k={order}", flush=True)
        oc, om, _ = O.search(od, order, 60, threads=2, mi_mode=2)
        for algorithm in (fa.ALGO_AUTO, fa.ALGO_MARGINAL) + ((fa.ALGO_DIRECT,) if order < 5 else ()):
            combos, mi = fa.GpuSearch(algorithm=algorithm).run_arrays(d, order, fa.Distribution(40, order - 1), 60)
            assert combos.tolist() == oc.tolist() and np.array_equal(mi, om), (order, algorithm)


def test_small_chunks_and_candidate_overflow():
    # tiny chunks and a candidate buffer smaller than a chunk's survivors
    # exercise the threshold merge and the re-scan of overflowed chunks
    print("case: overflow / redo", flush=True)
    d, od = synth(90, 256, 256, 71)
    ref_c, ref_m = fa.GpuSearch().run_arrays(d, 2, fa.Distribution(90, 1), 64)
    oc, om, _ = O.search(od, 2, 64, threads=2, mi_mode=2)
    assert ref_c.tolist() == oc.tolist() and np.array_equal(ref_m, om)
    # without the sampled starting threshold the first chunks overflow
    s = fa.GpuSearch(chunk_combos=300, candidate_cap=128, sample_threshold=False)
    c, m = s.run_arrays(d, 2, fa.Distribution(90, 1), 64)
    assert c.tolist() == ref_c.tolist() and np.array_equal(m, ref_m)
    assert s.last_stats.redo_chunks > 0
    # with it, the same tiny buffer suffices
    s = fa.GpuSearch(chunk_combos=300, candidate_cap=128)
    c, m = s.run_arrays(d, 2, fa.Distribution(90, 1), 64)
    assert c.tolist() == ref_c.tolist() and np.array_equal(m, ref_m)


@pytest.mark.parametrize("order", [2, 3])
@pytest.mark.parametrize("where", ["first", "last"])
def test_sampled_threshold_with_planted_signal(order, where):
    # the threshold sample is the tail of the distribution (at most one
    # chunk of combinations): with the strongest combinations in the first
    # prefixes (outside the sample) or in the last ones (inside it), the
    # starting threshold must cut none of them
    print(f"case: planted {where} k={order}", flush=True)
    n, half = 40, (300 if order == 2 else 1500)   # enough samples for the planted one to rank first
    planted = tuple(range(order)) if where == "first" else tuple(range(n - order, n))
    spec = fa.synth_spec(n, half, half, seed=81, planted=planted)
    words, geno, cw, tw = fa.synth_generate(spec, words=True, geno=True)
    d = fa.Dataset.from_words(words, n, cw, tw, half, half)
    od = O.pack(geno, np.r_[np.zeros(half, np.uint8), np.ones(half, np.uint8)])
    oc, om, _ = O.search(od, order, 30, threads=2, mi_mode=2)
    assert oc[0].tolist() == list(planted)
    algos = (fa.ALGO_AUTO, fa.ALGO_MFMA) if order == 2 else (fa.ALGO_AUTO,)
    for algorithm in algos:
        for chunk in (0, 50):
            c, m = fa.GpuSearch(algorithm=algorithm, chunk_combos=chunk).run_arrays(
                d, order, fa.Distribution(n, order - 1), 30)
            assert c.tolist() == oc.tolist() and np.array_equal(m, om), (algorithm, chunk)


@pytest.mark.parametrize("order,snps,outputs,cap", [
    (2, 400, 100, 1 << 17),    # 79,800 candidates in one merge: histogram filter, then one sort
    (2, 400, 3000, 1 << 17),   # the filter keeps 2N = 6,000 entries
    (3, 60, 50, 1 << 16),      # 34,220 candidates
    (2, 300, 64, 6000),        # 6,000 candidates per merge, overflowing chunks re-scanned
])
def test_merge_of_many_candidates(order, snps, outputs, cap):
    # with no sampled starting threshold the first chunk pushes every
    # combination; the merge must still return exactly the oracle's top-N
    print(f"case: many candidates k={order} n={snps} N={outputs} cap={cap}", flush=True)
    d, od = synth(snps, 400, 380, 500 + snps)
    oc, om, _ = O.search(od, order, outputs, threads=4, mi_mode=2)
    s = fa.GpuSearch(sample_threshold=False, candidate_cap=cap, chunk_combos=1 << 20)
    c, m = s.run_arrays(d, order, fa.Distribution(snps, order - 1), outputs)
    assert c.tolist() == oc.tolist() and np.array_equal(m, om)
    if cap < 10000:
        assert s.last_stats.redo_chunks > 0

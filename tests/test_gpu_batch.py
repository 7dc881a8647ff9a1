"""GPU: batched searches (fiu_search_batch) and the scan server (fiu_server_*).

Every search run in a batch or through the server must return exactly what
it returns alone: the same combinations in the same order with bit-identical
float MI (counts are exact, the MI is computed in FP64 from the counts and
ties are broken by the combination, so no tolerance is needed).  The plain
single-search path is itself checked against the oracle in
test_gpu_parity.py; small cases here are also checked against the oracle
directly.
"""
import math

import numpy as np
import pytest

import fiuncho_amd as fa
import oracle as O

pytestmark = pytest.mark.gpu


def synth(snps, cases, ctrls, seed, planted=(), geno=False):
    spec = fa.synth_spec(snps, cases, ctrls, seed=seed, planted=planted)
    words, g, cw, tw = fa.synth_generate(spec, words=True, geno=geno)
    return words, g, cw, tw


def dataset(snps, cases, ctrls, seed, planted=()):
    words, _, cw, tw = synth(snps, cases, ctrls, seed, planted)
    return fa.Dataset.from_words(words, snps, cw, tw, cases, ctrls)


def incomplete(snps, cases, ctrls, seed):
    """Genotype 3 (outside {0,1,2}) for 5% of the calls: the direct kinds."""
    rng = np.random.default_rng(seed)
    geno = rng.integers(0, 3, size=(snps, cases + ctrls)).astype(np.int8)
    geno[rng.random(geno.shape) < 0.05] = 3
    od = O.pack(geno, np.r_[np.zeros(cases, np.uint8), np.ones(ctrls, np.uint8)])
    return fa.Dataset.from_words(od.words, snps, od.cases_words, od.ctrls_words, cases, ctrls), od


def plain(d, order, dist, outputs):
    return fa.GpuSearch().run_arrays(d, order, dist, outputs)


def same(a, b):
    return a[0].tolist() == b[0].tolist() and np.array_equal(a[1], b[1])


@pytest.fixture(scope="module")
def mixed():
    """Datasets of different sizes, class balance and kinds."""
    print("case: batch datasets", flush=True)
    sets = [
        (dataset(300, 700, 500, 101), 2, 50),        # order-2 marginal
        (dataset(160, 1500, 900, 102), 3, 40),       # order-3 marginal (pair table)
        (dataset(45, 600, 650, 103), 4, 30),         # order-4 marginal (fourth launch)
        (dataset(20, 400, 380, 107), 5, 30),         # order-5 marginal (fourth launch)
        (dataset(500, 130, 370, 104), 2, 100),       # ragged classes
        (dataset(90, 2100, 1900, 105, planted=(3, 50, 77)), 3, 20),
    ]
    d_inc, od_inc = incomplete(60, 300, 260, 106)
    sets.append(((d_inc), 3, 25))                   # order-3 direct (incomplete genotypes)
    sets.append((d_inc, 2, 25))                     # order-2 direct
    sets.append((d_inc, 4, 25))                     # order-4 direct (second launch)
    return sets, od_inc


def test_batch_matches_plain(mixed):
    sets, od_inc = mixed
    scans = [(d, k, fa.Distribution(d.snps, k - 1), n) for d, k, n in sets]
    print("case: batch of 9 mixed searches", flush=True)
    res, st = fa.search_batch(scans)
    assert st.steps > 0 and st.device_ms > 0
    for (d, k, n), r in zip(sets, res):
        assert same(r, plain(d, k, fa.Distribution(d.snps, k - 1), n)), (d.snps, k)
    # the incomplete-genotype searches against the oracle as well
    for k, r in ((3, res[6]), (2, res[7]), (4, res[8])):
        oc, om, _ = O.search(od_inc, k, 25, threads=2, mi_mode=2)
        assert r[0].tolist() == oc.tolist() and np.array_equal(r[1], om)


def test_batch_planted_is_first(mixed):
    sets, _ = mixed
    d, k, n = sets[5]
    res, _ = fa.search_batch([(d, k, fa.Distribution(d.snps, k - 1), n)])
    assert res[0][0][0].tolist() == [3, 50, 77]


def test_batch_distributions_and_repeats(mixed):
    # the same dataset several times with different distribution shares, and
    # with the step of a small batch (many steps per search)
    sets, _ = mixed
    d, k = sets[1][0], 3
    scans = [(d, k, fa.Distribution(d.snps, 2).layer(5, o), 30) for o in range(5)]
    cfg = fa.batch_config(step_ms=0.3)
    res, st = fa.search_batch(scans, cfg)
    assert st.steps >= 4   # slices at most double: 1, 1, 2, 4, ... first slices
    merged = {}
    for (dd, kk, dist, n), r in zip(scans, res):
        assert same(r, plain(dd, kk, dist, n))
        merged.update({tuple(c): float(v) for c, v in zip(r[0].tolist(), r[1])})
    full = plain(d, k, fa.Distribution(d.snps, 2), 30)
    top = sorted(merged.items(), key=lambda t: (-t[1], t[0]))[:30]
    assert [list(c) for c, _ in top] == full[0].tolist()


def test_batch_candidate_overflow_rescans(mixed):
    # a buffer far smaller than a first slice: ranges overflow and are scanned
    # again, results unchanged
    sets, _ = mixed
    d, k, n = sets[0]
    cfg = fa.batch_config(candidate_cap=1024)
    res, st = fa.search_batch([(d, k, fa.Distribution(d.snps, 1), n)], cfg)
    assert st.redo_ranges > 0
    assert same(res[0], plain(d, k, fa.Distribution(d.snps, 1), n))


def test_batch_small_cases_vs_oracle():
    print("case: batch small cases vs oracle", flush=True)
    cases = []
    for i, (snps, ca, ct, k, n) in enumerate([(40, 90, 70, 2, 30), (25, 64, 200, 3, 40), (14, 100, 33, 4, 1001),
                                              (12, 120, 90, 5, 800)]):
        words, geno, cw, tw = synth(snps, ca, ct, 200 + i, geno=True)
        d = fa.Dataset.from_words(words, snps, cw, tw, ca, ct)
        od = O.pack(geno, np.r_[np.zeros(ca, np.uint8), np.ones(ct, np.uint8)])
        cases.append((d, od, k, n))
    res, _ = fa.search_batch([(d, k, fa.Distribution(d.snps, k - 1), n) for d, _, k, n in cases])
    for (d, od, k, n), r in zip(cases, res):
        oc, om, _ = O.search(od, k, n, threads=2, mi_mode=2)
        assert r[0].tolist() == oc.tolist() and np.array_equal(r[1], om), (d.snps, k)
        assert len(r[1]) == min(n, math.comb(d.snps, k))


def test_batch_edge_cases():
    d = dataset(12, 70, 50, 41)
    small = dataset(3, 30, 30, 42)
    scans = [(d, 3, fa.Distribution(12, 2), 0),          # no outputs
             (small, 4, fa.Distribution(3, 3), 10),      # fewer SNPs than the order
             (d, 2, fa.Distribution(12, 1, 5, 40), 10),  # offset past the last prefix
             (d, 2, fa.Distribution(12, 1), 1000)]       # every pair
    res, _ = fa.search_batch(scans)
    assert [len(r[1]) for r in res] == [0, 0, 0, math.comb(12, 2)]
    assert same(res[3], plain(d, 2, fa.Distribution(12, 1), 1000))
    with pytest.raises(fa.FiuError):
        fa.search_batch([(d, 6, fa.Distribution(12, 5), 10)])   # above order 5
    # order 5 on incomplete genotypes: every cell counted (its own kind)
    d_inc, od_inc = incomplete(12, 60, 60, 43)
    res, _ = fa.search_batch([(d_inc, 5, fa.Distribution(12, 4), 10), (d, 2, fa.Distribution(12, 1), 5)])
    assert same(res[0], plain(d_inc, 5, fa.Distribution(12, 4), 10))
    oc, om, _ = O.search(od_inc, 5, 10, threads=2, mi_mode=2)
    assert res[0][0].tolist() == oc.tolist() and np.array_equal(res[0][1], om)
    assert fa.search_batch([])[0] == []


def test_server_alone_and_in_the_mix():
    # each request alone (submit, wait) and all of them at once: identical
    # results, equal to the plain search on the same tables
    print("case: server alone vs mixed", flush=True)
    reqs = []
    for i, (snps, ca, ct, k, n) in enumerate([(400, 900, 1100, 2, 50), (150, 2000, 1000, 3, 30),
                                              (40, 500, 500, 4, 20), (700, 300, 300, 2, 80), (16, 700, 600, 5, 25),
                                              (120, 800, 1600, 3, 60), (64, 4000, 4000, 3, 10)]):
        words, _, cw, tw = synth(snps, ca, ct, 300 + i)
        reqs.append(dict(order=k, outputs=n, words=words, snps=snps, cases_words=cw, ctrls_words=tw,
                         cases=ca, ctrls=ct))
    srv = fa.Server()
    alone = []
    for r in reqs:
        c, m, st = srv.wait(srv.submit(**r))
        alone.append((c, m))
        assert st.latency_ms >= 0 and st.steps >= 1
    ids = [srv.submit(**r) for r in reqs]
    mixed = [srv.wait(i)[:2] for i in ids]
    info = srv.info()
    srv.close()
    for r, a, b in zip(reqs, alone, mixed):
        assert same(a, b)
        d = fa.Dataset.from_words(r["words"], r["snps"], r["cases_words"], r["ctrls_words"], r["cases"],
                                  r["ctrls"])
        assert same(a, plain(d, r["order"], fa.Distribution(r["snps"], r["order"] - 1), r["outputs"]))
    assert info.completed == 2 * len(reqs) and info.memory_used == 0


def test_server_admission_by_footprint():
    # a budget that holds about one request at a time: the rest wait in the
    # queue and are admitted as memory frees; a request that can never fit is
    # refused
    print("case: server admission", flush=True)
    words, _, cw, tw = synth(200, 1000, 1000, 401)
    # one request: 196,608 B of device tables + 153,600 B of upload staging +
    # row counts and binomials (about 362 kB); resident after the upload:
    # about 208 kB, so a 550 kB budget never admits two at once
    srv = fa.Server(fa.batch_config(memory_budget=550_000))
    kw = dict(order=2, outputs=40, words=words, snps=200, cases_words=cw, ctrls_words=tw, cases=1000, ctrls=1000)
    ids = [srv.submit(**kw) for _ in range(6)]
    res = [srv.wait(i) for i in ids]
    d = fa.Dataset.from_words(words, 200, cw, tw, 1000, 1000)
    ref = plain(d, 2, fa.Distribution(200, 1), 40)
    for c, m, st in res:
        assert same((c, m), ref)
        assert st.admit_ms >= st.submit_ms
    # later requests were admitted after earlier ones finished
    assert res[-1][2].admit_ms >= min(r[2].done_ms for r in res[:2])
    big, _, bcw, btw = synth(3000, 1000, 1000, 402)
    with pytest.raises(fa.FiuError) as e:
        srv.submit(order=2, outputs=10, words=big, snps=3000, cases_words=bcw, ctrls_words=btw, cases=1000,
                   ctrls=1000)
    assert e.value.status == fa._lib.FIU_ERR_REJECTED
    srv.close()


def test_server_resident_dataset_and_poll():
    d = dataset(250, 1200, 800, 501)
    srv = fa.Server()
    rid = srv.submit(order=3, outputs=25, dataset=d)
    out = srv.wait(rid, timeout_ms=0.0)
    while out is None:
        out = srv.wait(rid, timeout_ms=50.0)
    assert same(out[:2], plain(d, 3, fa.Distribution(250, 2), 25))
    assert out[2].combinations == math.comb(250, 3)
    assert out[2].algorithm == fa.ALGO_MFMA   # whole order-3 scans run on the matrix cores
    with pytest.raises(fa.FiuError):
        srv.submit(order=9, outputs=5, dataset=d)
    srv.close()

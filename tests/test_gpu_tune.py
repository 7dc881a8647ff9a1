"""GPU: kernel variants and the autotuner.

Every variant the tuner can pick (matrix-core staging depth / operand
format, popcount prefixes per wavefront) returns exactly the oracle's top-N,
and a tuned table steers AUTO searches and batched scans without changing
any result."""
import numpy as np
import pytest

import fiuncho_amd as fa
import oracle as O
from fiuncho_amd import tune

pytestmark = pytest.mark.gpu


def synth(snps, cases, ctrls, seed):
    spec = fa.synth_spec(snps, cases, ctrls, seed=seed)
    words, geno, cw, tw = fa.synth_generate(spec, words=True, geno=True)
    d = fa.Dataset.from_words(words, snps, cw, tw, cases, ctrls)
    od = O.pack(geno, np.r_[np.zeros(cases, np.uint8), np.ones(ctrls, np.uint8)])
    return d, od


@pytest.fixture(autouse=True)
def empty_table():
    tune.clear()
    yield
    tune.clear()


@pytest.mark.parametrize("order,snps,cases,ctrls", [(2, 150, 700, 500), (3, 70, 900, 1100)])
def test_every_variant_matches_the_oracle(order, snps, cases, ctrls):
    print(f"case: variants k={order} {snps}x{cases}+{ctrls}", flush=True)
    d, od = synth(snps, cases, ctrls, 71 + order)
    oc, om, _ = O.search(od, order, 40, threads=4, mi_mode=2)
    for algo, var in tune.candidates(order):
        s = fa.GpuSearch(algorithm=algo, variant=var)
        c, m = s.run_arrays(d, order, fa.Distribution(snps, order - 1), 40)
        assert s.last_stats.algorithm == algo and s.last_stats.variant == var, (algo, var)
        assert c.tolist() == oc.tolist() and np.array_equal(m, om), (algo, var)


def test_unknown_variant_is_refused():
    d, _ = synth(40, 100, 100, 5)
    with pytest.raises(fa.FiuError):
        fa.GpuSearch(algorithm=fa.ALGO_MARGINAL, variant=3).run_arrays(d, 2, fa.Distribution(40, 1), 5)
    with pytest.raises(fa.FiuError):
        fa.GpuSearch(algorithm=fa.ALGO_MFMA, variant=8).run_arrays(d, 3, fa.Distribution(40, 2), 5)


def test_autotune_steers_auto_searches_and_batches():
    classes = [tune.ShapeClass(2, 0, 1000, 0, 100000, 200, 2000),
               tune.ShapeClass(3, 0, 1000, 0, 100000, 90, 2000)]
    print("case: autotune two small classes", flush=True)
    entries = tune.autotune(classes, outputs=20, repeats=1)
    assert [e.order for e in entries] == [2, 3]
    assert [(e.algorithm, e.variant) for e in tune.get_table()] == [(e.algorithm, e.variant) for e in entries]
    for e in entries:
        assert (e.algorithm, e.variant) in tune.candidates(e.order) and e.ms > 0
    for k, snps in ((2, 300), (3, 120)):
        d, od = synth(snps, 800, 900, 90 + k)
        s = fa.GpuSearch()
        c, m = s.run_arrays(d, k, fa.Distribution(snps, k - 1), 30)
        e = entries[k - 2]
        assert (s.last_stats.algorithm, s.last_stats.variant) == (e.algorithm, e.variant)
        oc, om, _ = O.search(od, k, 30, threads=4, mi_mode=2)
        assert c.tolist() == oc.tolist() and np.array_equal(m, om)
        res, _ = fa.search_batch([(d, k, fa.Distribution(snps, k - 1), 30)])
        assert res[0][0].tolist() == oc.tolist() and np.array_equal(res[0][1], om)
    # a class whose tuned algorithm the dataset cannot use falls back
    tune.set_table([tune.Entry(2, 0, 1000, 0, 100000, fa.ALGO_MARGINAL, 8)])
    geno = np.random.default_rng(3).integers(0, 3, size=(30, 200)).astype(np.int8)
    geno[0, 5] = 3   # one uncalled sample: not every SNP covers the same samples
    od = O.pack(geno, (np.arange(200) >= 100).astype(np.uint8))
    d = fa.Dataset.from_words(od.words, 30, od.cases_words, od.ctrls_words, 100, 100)
    s = fa.GpuSearch()
    c, m = s.run_arrays(d, 2, fa.Distribution(30, 1), 10)
    assert s.last_stats.algorithm == fa.ALGO_DIRECT and s.last_stats.variant == 0
    oc, om, _ = O.search(od, 2, 10, threads=2, mi_mode=2)
    assert c.tolist() == oc.tolist() and np.array_equal(m, om)


def test_batch_unit_sizes_keep_results():
    # the batched popcount kinds' unit size (a tuning entry's batch_unit)
    # changes how the work is cut, not what is found
    from fiuncho_amd import tune
    print("case: batch unit sizes, order 4", flush=True)
    dss = [synth(snps, 900, 700, 140 + snps)[0] for snps in (60, 75, 90)]
    scans = [(d, 4, fa.Distribution(d.snps, 3), 40) for d in dss]
    plain = [fa.GpuSearch().run_arrays(*sc) for sc in scans]
    try:
        for unit in (1, 4, 6):
            tune.set_table([tune.Entry(4, 0, 1000, 0, 100000, fa.ALGO_MARGINAL, 0, 0.0, unit)])
            got, _ = fa.search_batch(scans)
            for (c, m), (pc, pm) in zip(got, plain):
                assert c.tolist() == pc.tolist() and np.array_equal(m, pm), unit
        # the tuner times every unit size of a class and keeps one
        table = tune.autotune_batch([tune.BatchClass(4, 0, 1000, 0, 100000, 70, 1600, count=2)], units=(0, 2, 5),
                                    repeats=1)
        entry = [e for e in table if e.order == 4]
        assert len(entry) == 1 and entry[0].batch_unit in (0, 2, 5)
        assert entry[0].algorithm == fa.ALGO_MARGINAL
    finally:
        tune.clear()
        for d in dss:
            d.close()


def test_batch_tiles_keep_results():
    # a matrix-core entry's batch tile (batch_unit 1 = narrow, 2 = wide)
    # changes which tile the batched scans run, not what they find; small
    # datasets (odd tile counts, ragged classes) take the wide tiles only this
    # way
    print("case: batch tiles, orders 2 and 3", flush=True)
    sets = [synth(300, 900, 700, 150), synth(130, 1000, 1100, 151), synth(700, 800, 900, 152),
            synth(517, 300, 1300, 153)]
    scans = [(sets[0][0], 3, fa.Distribution(300, 2), 40), (sets[1][0], 3, fa.Distribution(130, 2), 40),
             (sets[2][0], 2, fa.Distribution(700, 1), 60), (sets[3][0], 2, fa.Distribution(517, 1), 60)]
    want = [fa.GpuSearch().run_arrays(*sc) for sc in scans]
    for k, (d, od) in enumerate(sets[:2]):
        oc, om, _ = O.search(od, 3, 40, threads=4, mi_mode=2)
        assert want[k][0].tolist() == oc.tolist() and np.array_equal(want[k][1], om)
    E = tune.Entry
    try:
        for tile in (1, 2):
            tune.set_table([E(3, 0, 100000, 0, 1 << 20, fa.ALGO_MFMA, 0, 0.0, tile),
                            E(2, 0, 100000, 0, 1 << 20, fa.ALGO_MFMA, 0, 0.0, tile)])
            got, _ = fa.search_batch(scans)
            for (c, m), (wc, wm), sc in zip(got, want, scans):
                assert c.tolist() == wc.tolist() and np.array_equal(m, wm), (tile, sc[0].snps, sc[1])
        # the tuner times both tiles of a matrix-core class and keeps one
        tune.clear()
        table = tune.autotune_batch([tune.BatchClass(3, 0, 1000, 0, 100000, 300, 1600, count=2,
                                                     algorithm=fa.ALGO_MFMA)], repeats=1)
        entry = [e for e in table if e.order == 3]
        assert len(entry) == 1 and entry[0].algorithm == fa.ALGO_MFMA and entry[0].batch_unit in (1, 2)
    finally:
        tune.clear()
        for d, _ in sets:
            d.close()

"""GPU: orders 4 to 6, marginal in the last SNP (k_scan_km), and orders 7
to 9, every cell counted (k_scan_any).

The prefix's 3^(k-1) genotype rows against genotypes 0 and 1 of the last
SNP, genotype 2 from the prefix rows' own counts, the rows split over 3
(order 4), 9 (order 5) or 27 (order 6) lanes per last SNP.  Bar: the same
top-N (combinations, order and float MI bit for bit) as the oracle's FP64
ranking, as the direct all-cells kernel at order 4, and the reference's own
top-90 at orders 5 to 8 (tests/data, in test_gpu_parity) and its top lists
of synthetic sets at orders 7 to 9 (test_search_synthetic_golden)."""
import math

import numpy as np
import pytest

import fiuncho_amd as fa
import oracle as O

from .test_gpu_parity import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("order,snps,cases,ctrls,outputs,seed", [
    (4, 23, 333, 190, 100, 31),     # ragged classes, several c tiles of 21
    (4, 40, 64, 64, 500, 32),       # whole words only
    (4, 5, 50, 70, 5, 33),          # every combination of a tiny set
    (4, 4, 20, 30, 1, 34),          # the single combination
    (5, 15, 300, 260, 200, 35),     # tiles of 7 last SNPs
    (5, 12, 1000, 900, 792, 36),    # every combination (C(12, 5) = 792), 8 + 8 quads
    (5, 6, 40, 2000, 6, 37),        # 1 quad of cases, 16 of controls
    (5, 5, 70, 60, 1, 38),          # the single combination
    (6, 11, 300, 260, 200, 39),     # tiles of 2 last SNPs
    (6, 9, 1000, 900, 84, 40),      # every combination (C(9, 6) = 84)
    (6, 6, 50, 70, 1, 41),          # the single combination
])
def test_marginal_orders_vs_oracle(order, snps, cases, ctrls, outputs, seed):
    print(f"case: order {order} {snps}x{cases}+{ctrls}", flush=True)
    d, od = synth(snps, cases, ctrls, seed)
    oc, om, _ = O.search(od, order, outputs, threads=4, mi_mode=2)
    s = fa.GpuSearch(algorithm=fa.ALGO_MARGINAL)
    c, m = s.run_arrays(d, order, fa.Distribution(snps, order - 1), outputs)
    assert s.last_stats.algorithm == fa.ALGO_MARGINAL
    assert c.tolist() == oc.tolist() and np.array_equal(m, om)
    assert s.last_stats.combinations == math.comb(snps, order)
    auto = fa.GpuSearch()
    ca, ma = auto.run_arrays(d, order, fa.Distribution(snps, order - 1), outputs)
    assert auto.last_stats.algorithm == fa.ALGO_MARGINAL
    assert ca.tolist() == oc.tolist() and np.array_equal(ma, om)
    if order == 4:
        cd, md = fa.GpuSearch(algorithm=fa.ALGO_DIRECT).run_arrays(d, 4, fa.Distribution(snps, 3), outputs)
        assert cd.tolist() == oc.tolist() and np.array_equal(md, om)


@pytest.mark.parametrize("order", [4, 5, 6])
def test_marginal_orders_shares_and_chunks(order):
    snps = {4: 30, 5: 18, 6: 13}[order]
    planted = {4: (3, 11, 20, 27), 5: (1, 6, 9, 13, 17), 6: (0, 2, 5, 7, 9, 12)}[order]
    print(f"case: order {order} shares, chunks", flush=True)
    d, od = synth(snps, 700, 650, 90 + order, planted=planted)
    full = fa.GpuSearch().run_arrays(d, order, fa.Distribution(snps, order - 1), 300)
    oc, om, _ = O.search(od, order, 300, threads=4, mi_mode=2)
    assert full[0].tolist() == oc.tolist() and np.array_equal(full[1], om)
    # round-robin shares of the prefixes (MPI ranks) merge into the full list.
    # A share is the distribution's own prefixes, as ThreadedSearch(1) scans
    # them: with T threads the reference refines it to layer(T, t), another
    # set for step > 1 (ThreadedSearch.h:219-239, Distribution.h:172-176)
    merged = {}
    for off in range(3):
        dist = fa.Distribution(snps, order - 1).layer(3, off)
        c, m = fa.GpuSearch().run_arrays(d, order, dist, 300)
        sc, sm, _ = O.search(od, order, 300, step=3, offset=off, threads=1, mi_mode=2)
        assert c.tolist() == sc.tolist() and np.array_equal(m, sm)
        merged.update({tuple(x): v for x, v in zip(c.tolist(), m)})
    best = sorted(merged.items(), key=lambda t: (-t[1], t[0]))[:300]
    assert [c for c, _ in best] == [tuple(c) for c in full[0].tolist()]
    # tiny chunks and a small candidate buffer: overflowing chunks are re-scanned
    s = fa.GpuSearch(chunk_combos=500, candidate_cap=128, sample_threshold=False)
    c, m = s.run_arrays(d, order, fa.Distribution(snps, order - 1), 60)
    oc, om, _ = O.search(od, order, 60, threads=4, mi_mode=2)
    assert c.tolist() == oc.tolist() and np.array_equal(m, om)
    assert s.last_stats.redo_chunks > 0


def test_incomplete_data():
    # one uncalled genotype: order 4 falls back to the direct kernel, orders
    # 5 and 6 count every cell with the prefix rows split over lanes, orders
    # 7 to 12 count every cell per workgroup -- all against the oracle
    rng = np.random.default_rng(7)
    geno = rng.integers(0, 3, size=(13, 300)).astype(np.int8)
    geno[4, 17] = 3
    is_ctrl = (np.arange(300) >= 140).astype(np.uint8)
    od = O.pack(geno, is_ctrl)
    d = fa.Dataset.from_words(od.words, 13, od.cases_words, od.ctrls_words, 140, 160)
    for order in (4, 5, 6, 7, 8, 9, 10, 11, 12):
        print(f"case: incomplete data order {order}", flush=True)
        s = fa.GpuSearch()
        c, m = s.run_arrays(d, order, fa.Distribution(13, order - 1), 40)
        assert s.last_stats.algorithm == fa.ALGO_DIRECT
        oc, om, _ = O.search(od, order, 40, threads=4, mi_mode=2)
        assert c.tolist() == oc.tolist() and np.array_equal(m, om), order
    with pytest.raises(fa.FiuError):
        fa.GpuSearch(algorithm=fa.ALGO_MARGINAL).run_arrays(d, 5, fa.Distribution(13, 4), 10)
    with pytest.raises(fa.FiuError):
        fa.GpuSearch().run_arrays(d, 13, fa.Distribution(13, 12), 10)   # above order 12


def test_orders_5_6_all_cells_on_complete_data():
    # FIU_ALGO_DIRECT at orders 5 and 6 (every cell counted) equals the
    # marginal scan and the oracle; shares and tiny chunks included
    for order, snps in ((5, 22), (6, 14)):
        print(f"case: order {order} all cells vs marginal", flush=True)
        d, od = synth(snps, 600, 520, 70 + order)
        dist = fa.Distribution(snps, order - 1)
        s = fa.GpuSearch(algorithm=fa.ALGO_DIRECT)
        c, m = s.run_arrays(d, order, dist, 200)
        assert s.last_stats.algorithm == fa.ALGO_DIRECT
        mc, mm = fa.GpuSearch(algorithm=fa.ALGO_MARGINAL).run_arrays(d, order, dist, 200)
        assert c.tolist() == mc.tolist() and np.array_equal(m, mm)
        oc, om, _ = O.search(od, order, 200, threads=4, mi_mode=2)
        assert c.tolist() == oc.tolist() and np.array_equal(m, om)
        share = fa.Distribution(snps, order - 1).layer(4, 3)
        c, m = fa.GpuSearch(algorithm=fa.ALGO_DIRECT).run_arrays(d, order, share, 50)
        oc, om, _ = O.search(od, order, 50, step=4, offset=3, threads=1, mi_mode=2)
        assert c.tolist() == oc.tolist() and np.array_equal(m, om)
        s = fa.GpuSearch(algorithm=fa.ALGO_DIRECT, chunk_combos=300, candidate_cap=64, sample_threshold=False)
        c, m = s.run_arrays(d, order, dist, 40)
        oc, om, _ = O.search(od, order, 40, threads=4, mi_mode=2)
        assert c.tolist() == oc.tolist() and np.array_equal(m, om)
        assert s.last_stats.redo_chunks > 0


@pytest.mark.parametrize("order,snps,cases,ctrls,outputs,seed", [
    (10, 12, 300, 260, 66, 81),     # every combination (C(12, 10) = 66)
    (10, 13, 130, 70, 100, 82),     # 286 combinations, ragged classes
    (11, 13, 200, 180, 78, 83),     # every combination (C(13, 11) = 78)
    (12, 14, 160, 150, 40, 84),     # 91 combinations
    (12, 12, 64, 64, 1, 85),        # the single combination
])
def test_orders_10_to_12_vs_oracle(order, snps, cases, ctrls, outputs, seed):
    # 3^k counts per class in global memory (k_scan_big)
    print(f"case: order {order} {snps}x{cases}+{ctrls}", flush=True)
    d, od = synth(snps, cases, ctrls, seed)
    oc, om, _ = O.search(od, order, outputs, threads=4, mi_mode=2)
    s = fa.GpuSearch()
    c, m = s.run_arrays(d, order, fa.Distribution(snps, order - 1), outputs)
    assert s.last_stats.algorithm == fa.ALGO_DIRECT
    assert c.tolist() == oc.tolist() and np.array_equal(m, om)
    assert s.last_stats.combinations == math.comb(snps, order)
    # a share and tiny chunks (re-scans)
    share = fa.Distribution(snps, order - 1).layer(3, 2)
    c, m = fa.GpuSearch().run_arrays(d, order, share, outputs)
    sc, sm, _ = O.search(od, order, outputs, step=3, offset=2, threads=1, mi_mode=2)
    assert c.tolist() == sc.tolist() and np.array_equal(m, sm)


def test_wide_keys_beyond_the_packed_limit():
    # combination keys are lexicographic ranks: orders 7 and 9 run on more
    # SNPs than k * ceil(log2 n) <= 64 allowed (512 and 128); a share of a
    # few prefixes keeps the scan small.  Every returned combination lies in
    # the share, and its MI is the FP64 MI of the oracle's counts of it.
    for order, snps, step in ((7, 600, (1 << 31) - 1), (9, 200, (1 << 31) - 1)):
        print(f"case: wide keys order {order}, {snps} SNPs", flush=True)
        d, od = synth(snps, 64, 64, 90 + order)
        total = math.comb(snps, order - 1)
        for off in (0, total // 2 % step, total % step - 1):
            dist = fa.Distribution(snps, order - 1, step, off)
            s = fa.GpuSearch()
            c, m = s.run_arrays(d, order, dist, 30)
            assert len(c) > 0
            assert (np.diff(c, axis=1) > 0).all() and c.max() < snps
            for x, v in zip(c.tolist(), m):
                # lexicographic rank of the prefix among the (order-1)-combinations
                rank, prev = 0, -1
                for i, xi in enumerate(x[:-1]):
                    for v2 in range(prev + 1, xi):
                        rank += math.comb(snps - 1 - v2, order - 2 - i)
                    prev = xi
                assert rank % step == off, (x, rank)
                cc, ct = O.counts(od, x)
                assert abs(float(v) - O.mi_f64(cc, ct, order, 64, 64)) <= 0.5 * float(np.spacing(np.float32(v))) + 1e-12
            keys = [(-float(v), tuple(x)) for x, v in zip(c.tolist(), m)]
            assert keys == sorted(keys)
            # the same share in tiny chunks: identical
            t = fa.GpuSearch(chunk_combos=200, candidate_cap=128, sample_threshold=False)
            c2, m2 = t.run_arrays(d, order, dist, 30)
            assert c2.tolist() == c.tolist() and np.array_equal(m2, m)


@pytest.mark.parametrize("order,snps,cases,ctrls,outputs,seed", [
    (7, 12, 300, 260, 200, 51),     # 792 combinations
    (7, 9, 130, 70, 36, 52),        # every combination (C(9, 7) = 36), ragged classes
    (7, 7, 50, 70, 1, 53),          # the single combination
    (8, 11, 300, 280, 165, 54),     # every combination (C(11, 8) = 165)
    (8, 10, 1000, 900, 30, 55),     # 8 + 8 quads per row
    (8, 8, 64, 64, 1, 56),          # the single combination, whole words
    (9, 11, 300, 260, 55, 58),      # every combination (C(11, 9) = 55)
    (9, 9, 64, 64, 1, 59),          # the single combination
])
def test_all_cells_orders_vs_oracle(order, snps, cases, ctrls, outputs, seed):
    # orders 7 to 9: every 3^k cell counted per combination (k_scan_any)
    print(f"case: order {order} {snps}x{cases}+{ctrls}", flush=True)
    d, od = synth(snps, cases, ctrls, seed)
    oc, om, _ = O.search(od, order, outputs, threads=4, mi_mode=2)
    s = fa.GpuSearch()
    c, m = s.run_arrays(d, order, fa.Distribution(snps, order - 1), outputs)
    assert s.last_stats.algorithm == fa.ALGO_DIRECT
    assert c.tolist() == oc.tolist() and np.array_equal(m, om)
    assert s.last_stats.combinations == math.comb(snps, order)
    with pytest.raises(fa.FiuError):
        fa.GpuSearch(algorithm=fa.ALGO_MARGINAL).run_arrays(d, order, fa.Distribution(snps, order - 1), 5)


def test_all_cells_order7_shares_and_chunks():
    snps = 13
    print("case: order 7 shares, chunks", flush=True)
    d, od = synth(snps, 400, 380, 57)
    full = fa.GpuSearch().run_arrays(d, 7, fa.Distribution(snps, 6), 300)
    oc, om, _ = O.search(od, 7, 300, threads=4, mi_mode=2)
    assert full[0].tolist() == oc.tolist() and np.array_equal(full[1], om)
    merged = {}
    for off in range(3):
        dist = fa.Distribution(snps, 6).layer(3, off)
        c, m = fa.GpuSearch().run_arrays(d, 7, dist, 300)
        sc, sm, _ = O.search(od, 7, 300, step=3, offset=off, threads=1, mi_mode=2)
        assert c.tolist() == sc.tolist() and np.array_equal(m, sm)
        merged.update({tuple(x): v for x, v in zip(c.tolist(), m)})
    best = sorted(merged.items(), key=lambda t: (-t[1], t[0]))[:300]
    assert [c for c, _ in best] == [tuple(c) for c in full[0].tolist()]
    s = fa.GpuSearch(chunk_combos=200, candidate_cap=64, sample_threshold=False)
    c, m = s.run_arrays(d, 7, fa.Distribution(snps, 6), 60)
    oc, om, _ = O.search(od, 7, 60, threads=4, mi_mode=2)
    assert c.tolist() == oc.tolist() and np.array_equal(m, om)
    assert s.last_stats.redo_chunks > 0

"""GPU: the order-3 matrix-core scan (FIU_ALGO_MFMA at order 3).

The eight counts with genotypes in {0,1} are fp4 products of b rows masked by
a row of the first SNP with c rows (fp32 accumulation, exact below 2^24
samples per class); the other cells come from the pair table, the singles
and the totals as in the popcount scan.  Bar: the same top-N (combinations,
order and float MI bit for bit) as the oracle's FP64 ranking and as the
popcount scan.
"""
import math

import numpy as np
import pytest

import fiuncho_amd as fa
import oracle as O

from .test_gpu_parity import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("snps,cases,ctrls,outputs,seed", [
    (50, 333, 190, 60, 23),      # one tile, ragged classes
    (130, 300, 420, 200, 61),    # three tiles, partial last tile
    (23, 64, 64, 1771, 62),      # every triple
    (90, 1000, 900, 100, 63),    # 8 + 8 quads per class
    (3, 50, 60, 5, 64),          # the single triple
    (5, 64, 64, 10, 65),         # every triple of a tiny set
    (64, 200, 200, 300, 66),     # exactly one full tile
    (65, 40, 2000, 500, 67),     # one SNP in the second tile; 1 quad of cases, 16 of controls
])
def test_mfma3_vs_oracle(snps, cases, ctrls, outputs, seed):
    print(f"case: mfma3 {snps}x{cases}+{ctrls}", flush=True)
    d, od = synth(snps, cases, ctrls, seed)
    oc, om, _ = O.search(od, 3, outputs, threads=4, mi_mode=2)
    # default (64 x 32, 4-stage ring, at these sizes), 64 x 32 with 3 stages, 64 x 64
    # with 3 and 4 stages
    for variant in (0, 3, 5, 6):
        s = fa.GpuSearch(algorithm=fa.ALGO_MFMA, variant=variant)
        c, m = s.run_arrays(d, 3, fa.Distribution(snps, 2), outputs)
        assert s.last_stats.algorithm == fa.ALGO_MFMA
        assert c.tolist() == oc.tolist() and np.array_equal(m, om), variant
        assert s.last_stats.combinations == math.comb(snps, 3)


def test_mfma3_default_and_planted():
    # AUTO takes the matrix cores for whole order-3 scans; same result as the
    # popcount scan, planted triple first
    print("case: mfma3 400 SNPs planted", flush=True)
    d, _ = synth(400, 1500, 1500, 71, planted=(5, 200, 399))
    ref = fa.GpuSearch(algorithm=fa.ALGO_MARGINAL).run_arrays(d, 3, fa.Distribution(400, 2), 300)
    s = fa.GpuSearch()
    c, m = s.run_arrays(d, 3, fa.Distribution(400, 2), 300)
    assert s.last_stats.algorithm == fa.ALGO_MFMA
    assert c.tolist() == ref[0].tolist() and np.array_equal(m, ref[1])
    assert c[0].tolist() == [5, 200, 399]


def test_mfma3_distribution_and_chunks():
    # a round-robin share (masked in the epilogue) and tiny chunks with the
    # candidate re-scan; AUTO keeps the popcount scan for shares
    d, od = synth(40, 300, 280, 81)   # 9880 triples
    ref = {}
    full = fa.GpuSearch(algorithm=fa.ALGO_MARGINAL).run_arrays(d, 3, fa.Distribution(40, 2), 4096)
    for off in range(3):
        dist = fa.Distribution(40, 2).layer(3, off)
        s = fa.GpuSearch(algorithm=fa.ALGO_MFMA)
        c, m = s.run_arrays(d, 3, dist, 4096)
        p = fa.GpuSearch(algorithm=fa.ALGO_MARGINAL).run_arrays(d, 3, dist, 4096)
        assert c.tolist() == p[0].tolist() and np.array_equal(m, p[1])
        auto = fa.GpuSearch()
        auto.run_arrays(d, 3, dist, 10)
        assert auto.last_stats.algorithm == fa.ALGO_MARGINAL
        ref.update({tuple(cc): v for cc, v in zip(c.tolist(), m)})
    best = sorted(ref.items(), key=lambda t: (-t[1], t[0]))[:4096]
    assert [c for c, _ in best] == [tuple(c) for c in full[0].tolist()]
    s = fa.GpuSearch(algorithm=fa.ALGO_MFMA, chunk_combos=600, candidate_cap=256, sample_threshold=False)
    c, m = s.run_arrays(d, 3, fa.Distribution(40, 2), 64)
    oc, om, _ = O.search(od, 3, 64, threads=2, mi_mode=2)
    assert c.tolist() == oc.tolist() and np.array_equal(m, om)
    assert s.last_stats.redo_chunks > 0


@pytest.mark.parametrize("snps,variant", [(4200, 4), (5300, 6)])
def test_mfma3_launch_beyond_the_grid_limit(snps, variant):
    # one chunk of all first SNPs needs more than 2^32 work-items (4,200 SNPs,
    # 64 x 32 tiles: 553 x 4,198 x 8 = 18.6M workgroups of 256 threads;
    # 5,300 SNPs, 64 x 64 tiles: 436 x 5,298 x 8 = 18.5M): the launch is split
    # over first-SNP ranges, and the result equals the default chunks'
    print(f"case: mfma3 {snps} SNPs in one chunk, variant {variant}", flush=True)
    d, _ = synth(snps, 64, 64, 91)
    dist = fa.Distribution(snps, 2)
    ref = fa.GpuSearch(algorithm=fa.ALGO_MFMA, variant=variant).run_arrays(d, 3, dist, 50)
    s = fa.GpuSearch(algorithm=fa.ALGO_MFMA, variant=variant, chunk_combos=1 << 40, sample_threshold=False)
    c, m = s.run_arrays(d, 3, dist, 50)
    assert s.last_stats.combinations == math.comb(snps, 3)
    assert c.tolist() == ref[0].tolist() and np.array_equal(m, ref[1])

"""GPU: the order-2 matrix-core scan (FIU_ALGO_MFMA).

The marginal pair counts come from indicator dot products on the matrix cores
-- int8 operands with int32 accumulation, or fp4 (e2m1) operands with fp32
accumulation, both exact at these sizes; fp4 tiles of 128 x 128 or 64 x 128
pairs per workgroup (counts kept as 16-bit pairs between the classes,
two-pass epilogue) or 64 x 64; everything downstream is the popcount scan's.  Bar: the same
top-N (combinations, order and float MI bit for bit) as the oracle's FP64
ranking and as the popcount scan.
"""
import math

import numpy as np
import pytest

import fiuncho_amd as fa
import oracle as O

from .test_gpu_parity import synth

pytestmark = pytest.mark.gpu


# variant 0: the default (fp4, 64 x 64 pairs per workgroup at these sizes); 1:
# int8 operands; 2: fp4, 64 x 64; 3: fp4, 64 x 128; 4: fp4, 128 x 128
@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("snps,cases,ctrls,outputs,seed", [
    (70, 129, 63, 40, 21),       # one partial tile, ragged classes
    (128, 512, 512, 100, 22),    # two full tiles per side
    (200, 1000, 777, 300, 26),   # several tiles, K of 8 and 7 stages
    (65, 40, 2000, 2080, 27),    # every pair, one SNP in the last tile
    (150, 300, 280, 150, 28),    # 3 tiles: the last 128-SNP column half past the SNPs
    (530, 700, 650, 500, 29),    # 9 tiles: two super-blocks per side, odd last column
])
def test_mfma_vs_oracle(snps, cases, ctrls, outputs, seed, variant):
    print(f"case: mfma {snps}x{cases}+{ctrls} variant={variant}", flush=True)
    d, od = synth(snps, cases, ctrls, seed)
    oc, om, _ = O.search(od, 2, outputs, threads=4, mi_mode=2)
    s = fa.GpuSearch(algorithm=fa.ALGO_MFMA, variant=variant)
    c, m = s.run_arrays(d, 2, fa.Distribution(snps, 1), outputs)
    assert s.last_stats.algorithm == fa.ALGO_MFMA
    assert c.tolist() == oc.tolist() and np.array_equal(m, om)
    assert s.last_stats.combinations == math.comb(snps, 2)


def test_mfma_matches_popcount_scan():
    print("case: mfma vs popcount 1500 SNPs", flush=True)
    d, _ = synth(1500, 3000, 2500, 31, planted=(100, 1200))
    ref = fa.GpuSearch(algorithm=fa.ALGO_MARGINAL).run_arrays(d, 2, fa.Distribution(1500, 1), 500)
    for int8, variant in ((False, 0), (True, 0), (False, 2), (False, 3), (False, 4)):
        c, m = fa.GpuSearch(algorithm=fa.ALGO_MFMA, mfma_int8=int8, variant=variant).run_arrays(
            d, 2, fa.Distribution(1500, 1), 500)
        assert c.tolist() == ref[0].tolist() and np.array_equal(m, ref[1])
        assert c[0].tolist() == [100, 1200]


@pytest.mark.parametrize("variant", [0, 3, 4])
def test_mfma_distribution_and_chunks(variant):
    # distribution shares (masked in the epilogue) and tiny tile chunks with
    # the candidate re-scan, on the 64 x 64 (default here), 64 x 128 and
    # 128 x 128 tiles
    print(f"case: mfma shares and chunks variant={variant}", flush=True)
    d, od = synth(90, 300, 280, 41)   # 4005 pairs: every one fits in the top-4096
    full = fa.GpuSearch(algorithm=fa.ALGO_MARGINAL).run_arrays(d, 2, fa.Distribution(90, 1), 4096)
    ref = {tuple(c): v for c, v in zip(full[0].tolist(), full[1])}
    seen = {}
    for off in range(3):
        c, m = fa.GpuSearch(algorithm=fa.ALGO_MFMA, variant=variant).run_arrays(
            d, 2, fa.Distribution(90, 1).layer(3, off), 4096)
        for cc, mm in zip(c.tolist(), m):
            assert cc[0] % 3 == off and tuple(cc) not in seen
            seen[tuple(cc)] = mm
    assert seen == ref
    s = fa.GpuSearch(algorithm=fa.ALGO_MFMA, variant=variant, chunk_combos=4096 * 2, candidate_cap=256,
                     sample_threshold=False)
    c, m = s.run_arrays(d, 2, fa.Distribution(90, 1), 64)
    oc, om, _ = O.search(od, 2, 64, threads=2, mi_mode=2)
    assert c.tolist() == oc.tolist() and np.array_equal(m, om)
    assert s.last_stats.redo_chunks > 0


def test_mfma_refuses_what_it_cannot_do():
    d, _ = synth(40, 100, 100, 51)
    with pytest.raises(fa.FiuError):
        fa.GpuSearch(algorithm=fa.ALGO_MFMA).run(d, 4, fa.Distribution(40, 3), 5)
    rng = np.random.default_rng(52)
    geno = rng.integers(0, 3, size=(20, 150)).astype(np.int8)
    geno[0, 5] = 3   # a sample outside every row of SNP 0
    od = O.pack(geno, (np.arange(150) >= 80).astype(np.uint8))
    inc = fa.Dataset.from_words(od.words, 20, od.cases_words, od.ctrls_words, 80, 70)
    with pytest.raises(fa.FiuError):
        fa.GpuSearch(algorithm=fa.ALGO_MFMA).run(inc, 2, fa.Distribution(20, 1), 5)

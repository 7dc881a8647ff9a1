"""GPU: fiu_search_static, the search that allocates nothing.

It runs in the library's fixed per-GPU block instead of the dataset's
buffers; its results must equal fiu_search's on the same tables bit for bit
(combinations, order, float MI): counts are exact integers and the MI is the
FP64 value of the counts rounded, whatever kernel ran.  Calls follow each
other on the same block, larger inputs before smaller ones, so a stale
candidate list, top list, threshold or binomial table from the previous call
would show as a difference.
"""
import ctypes as C

import numpy as np
import pytest

import fiuncho_amd as fa
from fiuncho_amd import _lib

pytestmark = pytest.mark.gpu


def run_static(d, order, dist, outputs, cfg=None):
    cfg = cfg if cfg is not None else fa.Config()
    combos = np.zeros(max(outputs, 1) * order, np.int32)
    mi = np.zeros(max(outputs, 1), np.float32)
    cnt = C.c_uint32()
    stats = _lib.SearchStats()
    _lib.check(fa.lib().fiu_search_static(d.handle, order, dist.step, dist.offset, outputs, C.byref(cfg),
                                          combos.ctypes.data, mi.ctypes.data, C.byref(cnt), C.byref(stats)))
    n = cnt.value
    return combos[:n * order].reshape(n, order), mi[:n], stats


def words_of(snps, cases, ctrls, seed):
    spec = fa.synth_spec(snps, cases, ctrls, seed=seed)
    words, _, cw, tw = fa.synth_generate(spec, words=True)
    return words, cw, tw


def twin(snps, cases, ctrls, seed):
    """Two datasets on the same tables: one for each entry point, so that the
    static one starts without the tables fiu_search builds."""
    w, cw, tw = words_of(snps, cases, ctrls, seed)
    return (fa.Dataset.from_words(w, snps, cw, tw, cases, ctrls),
            fa.Dataset.from_words(w, snps, cw, tw, cases, ctrls))


def same(a, b):
    return a[0].tolist() == b[0].tolist() and np.array_equal(a[1], b[1])


def test_static_matches_fiu_search_on_calls_in_a_row():
    assert 0 < fa.lib().fiu_search_static_bytes() < 9 << 20
    cases = [
        # (snps, cases, ctrls, seed, order, outputs, distribution (step, offset))
        (400, 1500, 1200, 11, 3, 300, (1, 0)),    # larger first: order 3, no pair table yet -> direct kernel
        (90, 700, 650, 12, 2, 40, (1, 0)),        # smaller: order 2 (matrix cores, int8 operands)
        (150, 900, 1000, 13, 3, 25, (3, 1)),      # a distribution share
        (30, 800, 700, 14, 5, 30, (1, 0)),        # order 5
        (12, 400, 380, 15, 7, 20, (1, 0)),        # order 7
        (200, 333, 190, 16, 2, 4096, (1, 0)),     # every pair up to 4,096 outputs
        (64, 64, 64, 17, 4, 10, (1, 0)),          # smallest last
    ]
    for snps, ca, ct, seed, order, outputs, (step, off) in cases:
        print(f"case: static {snps}x{ca}+{ct} order {order} dist ({step},{off})", flush=True)
        d_static, d_plain = twin(snps, ca, ct, seed)
        dist = fa.Distribution(snps, order - 1, step, off)
        got = run_static(d_static, order, dist, outputs)
        plain = fa.GpuSearch()
        want = plain.run_arrays(d_plain, order, dist, outputs)
        assert same(got, want), (snps, order)
        assert got[2].combinations == plain.last_stats.combinations
        # the same dataset after fiu_search built its tables: the static
        # search uses them (matrix cores at orders 2-3), same results again
        fa.GpuSearch().run_arrays(d_static, order, dist, outputs)
        again = run_static(d_static, order, dist, outputs)
        assert same(again, want), (snps, order, "prepared")
        d_static.close()
        d_plain.close()


def test_static_incomplete_genotypes_and_tiny_chunks():
    rng = np.random.default_rng(21)
    geno = rng.integers(0, 3, size=(70, 900)).astype(np.int8)
    geno[rng.random(geno.shape) < 0.03] = 3   # uncalled
    is_ctrl = (np.arange(900) >= 480).astype(np.uint8)
    d = fa.Dataset.from_genotypes(geno, is_ctrl)
    for order, outputs in ((3, 50), (2, 30)):
        print(f"case: static incomplete order {order}", flush=True)
        dist = fa.Distribution(70, order - 1)
        want = fa.GpuSearch().run_arrays(d, order, dist, outputs)
        got = run_static(d, order, dist, outputs)
        assert same(got, want), order
        # many launches, a small candidate buffer: overflowing chunks are
        # scanned again with the final threshold
        small = fa.Config(chunk_combos=500, candidate_cap=64, sample_threshold=False)
        got = run_static(d, order, dist, outputs, small)
        assert same(got, want), (order, "tiny chunks")
        assert got[2].redo_chunks > 0
    d.close()


def test_static_refuses_beyond_its_limits():
    d = fa.Dataset.synth(fa.synth_spec(60, 100, 100, seed=31))
    with pytest.raises(fa.FiuError) as e:
        run_static(d, 2, fa.Distribution(60, 1), 10, fa.Config(candidate_cap=(1 << 18) + 1))
    assert e.value.status == _lib.FIU_ERR_UNSUPPORTED
    d.close()
    # order x (snps + 1) above 2^19 binomial entries
    big = fa.Dataset.synth(fa.synth_spec(200_000, 1, 1, seed=32))
    with pytest.raises(fa.FiuError) as e:
        run_static(big, 3, fa.Distribution(200_000, 2), 10)
    assert e.value.status == _lib.FIU_ERR_UNSUPPORTED and "binomial" in str(e.value)
    big.close()

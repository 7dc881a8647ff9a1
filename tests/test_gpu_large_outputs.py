"""GPU: more than 4,096 results per search.

The reference's MaxArray is sized by `outputs` (MaxArray.h:31-46,
ThreadedSearch.h:241-244).  Above 4,096 results the top-N merge runs in
global memory (a bitonic sort of the list and the new candidates, then the
first N unique entries); the lists must equal the oracle's FP64 ranking bit
for bit, with and without re-scanned chunks.
"""
import math

import numpy as np
import pytest

import fiuncho_amd as fa
import oracle as O

from .test_gpu_parity import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("order,snps,outputs,algo", [
    (2, 150, 10000, fa.ALGO_AUTO),        # 11,175 pairs (matrix cores)
    (2, 150, 11175, fa.ALGO_MARGINAL),    # every pair
    (3, 60, 20000, fa.ALGO_AUTO),         # 34,220 triples
    (3, 60, 20000, fa.ALGO_DIRECT),
    (4, 26, 12000, fa.ALGO_AUTO),         # 14,950 quads
])
def test_large_outputs_vs_oracle(order, snps, outputs, algo):
    print(f"case: order {order}, {snps} SNPs, top-{outputs}", flush=True)
    d, od = synth(snps, 400, 380, 300 + order)
    oc, om, _ = O.search(od, order, outputs, threads=4, mi_mode=2)
    assert len(oc) == min(outputs, math.comb(snps, order))
    c, m = fa.GpuSearch(algorithm=algo).run_arrays(d, order, fa.Distribution(snps, order - 1), outputs)
    assert c.tolist() == oc.tolist() and np.array_equal(m, om)
    # small chunks, no starting threshold
    s = fa.GpuSearch(algorithm=algo, chunk_combos=2000, sample_threshold=False)
    c, m = s.run_arrays(d, order, fa.Distribution(snps, order - 1), outputs)
    assert c.tolist() == oc.tolist() and np.array_equal(m, om)
    # one chunk, no threshold, the smallest candidate buffer (raised to
    # `outputs`): the chunk overflows and is scanned again
    s = fa.GpuSearch(algorithm=algo, candidate_cap=1, sample_threshold=False)
    c, m = s.run_arrays(d, order, fa.Distribution(snps, order - 1), outputs)
    assert c.tolist() == oc.tolist() and np.array_equal(m, om)
    assert (s.last_stats.redo_chunks > 0) == (math.comb(snps, order) > outputs)


def test_large_outputs_limits():
    d, _ = synth(40, 100, 100, 310)
    with pytest.raises(fa.FiuError):
        fa.GpuSearch().run_arrays(d, 2, fa.Distribution(40, 1), (1 << 24) + 1)
    # a share: the merged shares equal the whole list
    full = fa.GpuSearch().run_arrays(d, 3, fa.Distribution(40, 2), 9880)
    parts = {}
    for off in range(2):
        c, m = fa.GpuSearch().run_arrays(d, 3, fa.Distribution(40, 2).layer(2, off), 9880)
        parts.update({tuple(x): v for x, v in zip(c.tolist(), m)})
    best = sorted(parts.items(), key=lambda t: (-t[1], t[0]))
    assert [x for x, _ in best] == [tuple(x) for x in full[0].tolist()]

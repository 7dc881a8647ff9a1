"""Autotuner of the shape-class table (``fiu_set_tuning``).

For each shape class (order, range of SNP counts, range of sample counts) the
tuner builds a representative synthetic dataset on the device, times every
candidate kernel -- algorithm and variant: the matrix-core scans with their
LDS staging depth (order 3) or operand format (order 2), the popcount scans
with their prefixes per wavefront -- on the whole scan, requires every
candidate to return the identical top-N list, and keeps the fastest.  The
table is loaded into the library, where AUTO searches and batched scans of
that class use it.  ``autotune_batch`` then times ``fiu_search_batch`` over a
few datasets of each batch class with every unit size of the batched
popcount kernels, or both batched tiles of the matrix-core kernels (the
entry's ``batch_unit``), and keeps the fastest.

    entries = tune.autotune()          # times the default classes, loads the table
    entries = tune.autotune_batch()    # adds the batched kernels' unit sizes / tiles
    tune.save(entries, "tuning.json")  # later: tune.load("tuning.json")
"""
from __future__ import annotations

import ctypes as C
import json
import time
from dataclasses import asdict, dataclass
from typing import Iterable, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from ._lib import ALGO_DIRECT, ALGO_MARGINAL, ALGO_MFMA, FiuError, check, lib

INF = 2 ** 31 - 1


@dataclass
class Entry:
    """One class of the table: [snps_lo, snps_hi) x [samples_lo, samples_hi)."""
    order: int
    snps_lo: int
    snps_hi: int
    samples_lo: int
    samples_hi: int
    algorithm: int
    variant: int = 0
    ms: float = 0.0
    batch_unit: int = 0   # batched popcount kinds: 0 = default unit, k = default x 2^(k-3);
                          # matrix cores: batched tile, 0 = by size, 1 = narrow, 2 = wide


@dataclass
class ShapeClass:
    order: int
    snps_lo: int
    snps_hi: int
    samples_lo: int
    samples_hi: int
    snps: int       # representative dataset timed for the class
    samples: int


def _classes(order, snp_edges, snp_reps, sample_edges, sample_reps):
    out = []
    for (slo, shi), s in zip(zip(snp_edges[:-1], snp_edges[1:]), snp_reps):
        for (nlo, nhi), m in zip(zip(sample_edges[:-1], sample_edges[1:]), sample_reps):
            out.append(ShapeClass(order, slo, shi, nlo, nhi, s, m))
    return out


# Classes covering every order-2 and order-3 shape; each representative is
# timed on its whole scan (a few hundred milliseconds at most).
DEFAULT_CLASSES: List[ShapeClass] = (
    _classes(2, [0, 2048, 16384, INF], [1000, 6000, 20000], [0, 8192, INF], [4000, 16000])
    + _classes(3, [0, 512, 2048, INF], [300, 1000, 2000], [0, 8192, INF], [4000, 16000]))


def candidates(order: int) -> List[Tuple[int, int]]:
    """(algorithm, variant) pairs the tuner times at `order`."""
    if order == 2:
        return [(ALGO_MFMA, 1), (ALGO_MFMA, 2), (ALGO_MFMA, 3), (ALGO_MFMA, 4), (ALGO_MARGINAL, 2), (ALGO_MARGINAL, 4),
                (ALGO_MARGINAL, 8)]
    if order == 3:
        return [(ALGO_MFMA, 4), (ALGO_MFMA, 5), (ALGO_MFMA, 6), (ALGO_MARGINAL, 2), (ALGO_MARGINAL, 4)]
    return [(ALGO_DIRECT, 0)]


def set_table(entries: Sequence[Entry]) -> None:
    """Load `entries` as the library's table (an empty list clears it)."""
    arr = (_lib.TuneEntry * max(len(entries), 1))()
    for i, e in enumerate(entries):
        t = arr[i]
        t.order, t.snps_lo, t.snps_hi = e.order, e.snps_lo, e.snps_hi
        t.samples_lo, t.samples_hi = e.samples_lo, e.samples_hi
        t.algorithm, t.variant, t.ms, t.batch_unit = e.algorithm, e.variant, e.ms, e.batch_unit
    check(lib().fiu_set_tuning(arr, len(entries)))


def get_table() -> List[Entry]:
    n = C.c_size_t()
    check(lib().fiu_get_tuning(None, 0, C.byref(n)))
    arr = (_lib.TuneEntry * max(n.value, 1))()
    check(lib().fiu_get_tuning(arr, n.value, C.byref(n)))
    return [Entry(t.order, t.snps_lo, t.snps_hi, t.samples_lo, t.samples_hi, t.algorithm, t.variant, t.ms,
                  t.batch_unit) for t in arr[:n.value]]


def clear() -> None:
    set_table([])


def save(entries: Sequence[Entry], path: str) -> None:
    with open(path, "w") as f:
        json.dump([asdict(e) for e in entries], f, indent=1)


def load(path: str, apply: bool = True) -> List[Entry]:
    with open(path) as f:
        entries = [Entry(**e) for e in json.load(f)]
    if apply:
        set_table(entries)
    return entries


def _time(search, d, order, dist, outputs, repeats):
    best = float("inf")
    res = None
    for _ in range(repeats):
        t0 = time.perf_counter()
        res = search.run_arrays(d, order, dist, outputs)
        best = min(best, time.perf_counter() - t0)
    return best, res


def autotune(classes: Iterable[ShapeClass] = DEFAULT_CLASSES, outputs: int = 50, repeats: int = 2,
             seed: int = 7, apply: bool = True, log=None) -> List[Entry]:
    """Time every candidate of every class; returns the table (and loads it
    when `apply`).  Raises if two candidates disagree on a result."""
    from . import Dataset, Distribution, GpuSearch, synth_spec
    saved = get_table()
    clear()   # candidates are timed as requested, not through an older table
    entries: List[Entry] = []
    try:
        for sc in classes:
            cases = sc.samples // 2
            spec = synth_spec(sc.snps, cases, sc.samples - cases, seed=seed)
            d = Dataset.synth(spec)
            dist = Distribution(sc.snps, sc.order - 1)
            best: Optional[Tuple[float, int, int]] = None
            ref = None
            for algo, var in candidates(sc.order):
                s = GpuSearch(algorithm=algo, variant=var)
                try:
                    s.run_arrays(d, sc.order, dist, outputs)   # tables, code, caches
                except FiuError:
                    continue
                t, (c, m) = _time(s, d, sc.order, dist, outputs, repeats)
                if ref is None:
                    ref = (c, m)
                elif c.tolist() != ref[0].tolist() or not np.array_equal(m, ref[1]):
                    raise RuntimeError(f"candidate {algo}/{var} disagrees on class {sc}")
                if log:
                    log(f"[tune] order {sc.order} {sc.snps} x {sc.samples}: algorithm {algo} variant {var} "
                        f"{t * 1e3:.2f} ms")
                if best is None or t < best[0]:
                    best = (t, algo, var)
            d.close()
            if best is not None:
                entries.append(Entry(sc.order, sc.snps_lo, sc.snps_hi, sc.samples_lo, sc.samples_hi,
                                     best[1], best[2], best[0] * 1e3))
    except BaseException:
        set_table(saved)
        raise
    set_table(entries if apply else saved)
    return entries


@dataclass
class BatchClass:
    """A class of batched scans: [snps_lo, snps_hi) x [samples_lo, samples_hi)
    at `order`, timed on `count` datasets of `snps` x `samples` with
    `algorithm` (a popcount kind: the unit size has no effect on the
    matrix-core kinds)."""
    order: int
    snps_lo: int
    snps_hi: int
    samples_lo: int
    samples_hi: int
    snps: int
    samples: int
    count: int = 4
    algorithm: int = ALGO_MARGINAL


# Orders 4 and 5, the batched scans that run on the popcount kinds, and
# orders 2 and 3 at the sizes where the matrix cores' two tiles compete.
DEFAULT_BATCH_CLASSES: List[BatchClass] = [
    BatchClass(4, 0, 512, 0, 8192, 200, 3000),
    BatchClass(4, 0, 512, 8192, INF, 160, 16000),
    BatchClass(5, 0, 128, 0, INF, 50, 4000),
    BatchClass(3, 512, 2048, 8192, INF, 1000, 16000, count=4, algorithm=ALGO_MFMA),
    BatchClass(3, 2048, INF, 8192, INF, 2500, 16000, count=3, algorithm=ALGO_MFMA),
    BatchClass(2, 16384, INF, 8192, INF, 20000, 16000, count=3, algorithm=ALGO_MFMA),
]
BATCH_UNITS = (0, 1, 2, 4, 5, 6)
BATCH_TILES = (1, 2)   # matrix-core classes: narrow, wide


def autotune_batch(classes: Iterable[BatchClass] = DEFAULT_BATCH_CLASSES, units: Sequence[int] = BATCH_UNITS,
                   outputs: int = 50, repeats: int = 2, seed: int = 11, apply: bool = True, log=None) -> List[Entry]:
    """Time fiu_search_batch over each class's datasets with every unit size
    (popcount kinds) or both tiles (a class whose entry is on the matrix
    cores); the fastest goes into the class's table entry (added when the
    table has none for it).  Returns the whole table (loaded when `apply`).
    Raises if two candidates disagree on a result."""
    from . import Dataset, Distribution, search_batch, synth_spec
    saved = get_table()
    table = list(saved)
    try:
        for bc in classes:
            same = [e for e in table if (e.order, e.snps_lo, e.snps_hi, e.samples_lo, e.samples_hi) ==
                    (bc.order, bc.snps_lo, bc.snps_hi, bc.samples_lo, bc.samples_hi)]
            base = same[0] if same else Entry(bc.order, bc.snps_lo, bc.snps_hi, bc.samples_lo, bc.samples_hi,
                                              bc.algorithm)
            rest = [e for e in table if e not in same]
            dss = []
            for i in range(bc.count):
                cases = bc.samples // 2
                dss.append(Dataset.synth(synth_spec(bc.snps, cases, bc.samples - cases, seed=seed + i)))
            scans = [(d, bc.order, Distribution(bc.snps, bc.order - 1), outputs) for d in dss]
            best = None
            ref = None
            for u in (BATCH_TILES if base.algorithm == ALGO_MFMA else units):
                entry = Entry(base.order, base.snps_lo, base.snps_hi, base.samples_lo, base.samples_hi,
                              base.algorithm, base.variant, base.ms, u)
                set_table(rest + [entry])
                t = float("inf")
                res = None
                for _ in range(repeats):
                    t0 = time.perf_counter()
                    res, _ = search_batch(scans)
                    t = min(t, time.perf_counter() - t0)
                got = [(c.tolist(), m.tobytes()) for c, m, *_ in res]
                if ref is None:
                    ref = got
                elif got != ref:
                    raise RuntimeError(f"batch unit {u} disagrees on class {bc}")
                if log:
                    log(f"[tune] batch order {bc.order} {bc.count} x {bc.snps} x {bc.samples}: "
                        f"{'tile' if base.algorithm == ALGO_MFMA else 'unit'} {u} "
                        f"{t * 1e3:.2f} ms")
                if best is None or t < best[0]:
                    best = (t, u)
            for d in dss:
                d.close()
            table = rest + [Entry(base.order, base.snps_lo, base.snps_hi, base.samples_lo, base.samples_hi,
                                  base.algorithm, base.variant, base.ms, best[1])]
    except BaseException:
        set_table(saved)
        raise
    set_table(table if apply else saved)
    return table

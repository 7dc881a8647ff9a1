"""fiuncho_amd: MI355X-native exhaustive epistasis search.

Python mirror of the reference's search interface (UDC-GAC/fiuncho):

  ==========================  ==============================================
  this package                reference
  ==========================  ==============================================
  ``Dataset.read(paths)``     ``Dataset<uint64_t>::read`` (Dataset.h:48-94)
  ``Distribution``            ``Distribution<int>`` (Distribution.h:24-203)
  ``Result``                  ``Result<int,float>`` (utils/Result.h:17-70)
  ``GpuSearch(...).run``      ``Search::run`` (Search.h:40-43), in place of
                              ``ThreadedSearch::run`` (ThreadedSearch.h:207-246)
  ``combine_batch``           ``GenotypeTable::combine`` (GenotypeTable.h:123)
  ``contingency_batch``       ``GenotypeTable::combine_and_popcnt`` (:137-140)
  ``mi_batch``                ``MutualInformation<float>::compute`` (MI.h:83)
  ==========================  ==============================================

Every call goes through the C ABI of ``libfiuncho_amd.so`` (HIP kernels for
gfx950); nothing here computes on the CPU.
"""
from __future__ import annotations

import ctypes as C
import struct
from dataclasses import dataclass, field
from typing import List, Sequence

import numpy as np

from . import _lib
from ._lib import ALGO_AUTO, ALGO_DIRECT, ALGO_MARGINAL, ALGO_MFMA, FiuError, check, lib

__all__ = ["Dataset", "Distribution", "Result", "GpuSearch", "Config", "combine_batch",
           "contingency_batch", "mi_batch", "ct_size", "synth_spec", "synth_generate",
           "search_batch", "batch_config", "Server",
           "FiuError", "ALGO_AUTO", "ALGO_DIRECT", "ALGO_MARGINAL", "ALGO_MFMA", "version"]


def version() -> str:
    return lib().fiu_version().decode()


def ct_size(order: int) -> int:
    """Padded contingency-table length per class (ContingencyTable.h:53-65)."""
    return int(lib().fiu_ct_size(order))


# ---------------------------------------------------------------------------
# Synthetic data
# ---------------------------------------------------------------------------
def synth_spec(snps, cases, ctrls, seed=1, maf=(0.05, 0.5), low_frac=0.0, low=(0.01, 0.05),
               planted=(), penetrance_seed=2, penetrance=(0.3, 0.7), threads=0):
    sp = _lib.SynthSpec()
    lib().fiu_synth_spec_init(C.byref(sp))
    sp.snps, sp.cases, sp.ctrls, sp.seed = snps, cases, ctrls, seed
    sp.maf_lo, sp.maf_hi = maf
    sp.low_frac = low_frac
    sp.low_lo, sp.low_hi = low
    sp.planted_order = len(planted)
    for i, s in enumerate(planted):
        sp.planted[i] = s
    sp.penetrance_seed = penetrance_seed
    sp.pen_lo, sp.pen_hi = penetrance
    sp.threads = threads
    return sp


def read_files(inputs: Sequence[str]):
    """Host-only ``Dataset<uint64_t>::read``: (words, snps, cases, ctrls,
    cases_words, ctrls_words) in the reference host layout, no device."""
    arr = (C.c_char_p * len(inputs))(*[p.encode() for p in inputs])
    snps, cw, tw = C.c_size_t(), C.c_size_t(), C.c_size_t()
    cases, ctrls = C.c_uint32(), C.c_uint32()
    check(lib().fiu_read_files(arr, len(inputs), None, 0, C.byref(snps), C.byref(cases), C.byref(ctrls),
                               C.byref(cw), C.byref(tw)))
    words = np.zeros(max(snps.value * 3 * (cw.value + tw.value), 1), np.uint64)
    check(lib().fiu_read_files(arr, len(inputs), words.ctypes.data, words.size, None, None, None, None, None))
    n = snps.value * 3 * (cw.value + tw.value)
    return words[:n], snps.value, cases.value, ctrls.value, cw.value, tw.value


def synth_generate(spec, words=True, geno=False):
    """Host arrays (words in the reference layout, int8 genotypes)."""
    cw, tw = C.c_size_t(), C.c_size_t()
    check(lib().fiu_synth_generate(C.byref(spec), None, None, C.byref(cw), C.byref(tw)))
    n, N = spec.snps, spec.cases + spec.ctrls
    w = np.zeros(n * 3 * (cw.value + tw.value), np.uint64) if words else None
    g = np.zeros((n, N), np.int8) if geno else None
    check(lib().fiu_synth_generate(C.byref(spec), None if w is None else w.ctypes.data,
                                   None if g is None else g.ctypes.data, C.byref(cw), C.byref(tw)))
    return w, g, cw.value, tw.value


# ---------------------------------------------------------------------------
# Dataset
# ---------------------------------------------------------------------------
class Dataset:
    """Device-resident bit tables (reference layout: GenotypeTable.h:77-101)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)
        snps, cw, tw, dev = C.c_size_t(), C.c_size_t(), C.c_size_t(), C.c_size_t()
        cases, ctrls = C.c_uint32(), C.c_uint32()
        check(lib().fiu_dataset_info(self._h, C.byref(snps), C.byref(cases), C.byref(ctrls),
                                     C.byref(cw), C.byref(tw), C.byref(dev)))
        self.snps, self.cases, self.ctrls = snps.value, cases.value, ctrls.value
        self.cases_words, self.ctrls_words, self.device_bytes = cw.value, tw.value, dev.value

    @classmethod
    def read(cls, inputs: Sequence[str]) -> "Dataset":
        """``Dataset<uint64_t>::read``: one .raw, or a .tped/.tfam pair."""
        arr = (C.c_char_p * len(inputs))(*[p.encode() for p in inputs])
        h = C.c_void_p()
        check(lib().fiu_dataset_read(arr, len(inputs), C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_genotypes(cls, geno, is_ctrl, reader_flush: bool = False) -> "Dataset":
        """Pack a SNP-major genotype matrix (snps x samples; 0/1/2, other
        values uncalled) on the device; `is_ctrl` per sample.  `reader_flush`
        zeroes the words the reference file readers re-flush, as
        ``Dataset.read`` does (RAWFile.hpp:226-259)."""
        geno = np.ascontiguousarray(geno, dtype=np.int8)
        is_ctrl = np.ascontiguousarray(is_ctrl, dtype=np.uint8)
        if geno.ndim != 2 or is_ctrl.shape != (geno.shape[1],):
            raise ValueError("geno must be snps x samples and is_ctrl one flag per sample")
        h = C.c_void_p()
        check(lib().fiu_dataset_from_genotypes(geno.ctypes.data, geno.shape[0], geno.shape[1], is_ctrl.ctypes.data,
                                               _lib.FIU_GENO_READER_FLUSH if reader_flush else 0, C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_words(cls, words, snps, cases_words, ctrls_words, cases, ctrls) -> "Dataset":
        words = np.ascontiguousarray(words, dtype=np.uint64)
        if words.size != snps * 3 * (cases_words + ctrls_words):
            raise ValueError("words has the wrong length for the given shape")
        h = C.c_void_p()
        check(lib().fiu_dataset_create(words.ctypes.data, snps, cases_words, ctrls_words, cases,
                                       ctrls, C.byref(h)))
        return cls(h.value)

    @classmethod
    def synth(cls, spec) -> "Dataset":
        h = C.c_void_p()
        check(lib().fiu_dataset_synth(C.byref(spec), C.byref(h)))
        return cls(h.value)

    def words(self) -> np.ndarray:
        out = np.zeros(self.snps * 3 * (self.cases_words + self.ctrls_words), np.uint64)
        check(lib().fiu_dataset_words(self._h, out.ctypes.data))
        return out

    def close(self):
        if self._h and self._h.value:
            check(lib().fiu_dataset_destroy(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h


# ---------------------------------------------------------------------------
# Distribution / Result
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class Distribution:
    """Round-robin share of the (k-1)-prefixes (Distribution.h:24-203).

    Prefixes are enumerated in lexicographic order; this share takes those
    whose rank r satisfies r = offset + j*step."""
    n: int
    k: int
    step: int = 1
    offset: int = 0

    def layer(self, step: int, offset: int) -> "Distribution":
        """Distribution.h:172-176."""
        return Distribution(self.n, self.k, self.step * step, self.offset * step + offset)


@dataclass
class Result:
    """Result<int,float> (utils/Result.h:17-70)."""
    combination: List[int]
    val: float

    def str(self) -> str:
        """Result.h:37-45: indices then the value, space separated (C++
        ostream default float formatting, 6 significant digits)."""
        return " ".join(str(i) for i in self.combination) + " " + f"{self.val:.6g}"

    def serialize(self) -> bytes:
        """Result.h:47-55: size_t length, int32 indices, float value."""
        return struct.pack(f"<Q{len(self.combination)}if", len(self.combination),
                           *self.combination, self.val)

    @staticmethod
    def deserialize(buf: bytes, pos: int = 0):
        """Result.h:57-66; returns (Result, next position)."""
        (k,) = struct.unpack_from("<Q", buf, pos)
        vals = struct.unpack_from(f"<{k}if", buf, pos + 8)
        return Result(list(vals[:k]), vals[k]), pos + 8 + 4 * k + 4


# ---------------------------------------------------------------------------
# Search
# ---------------------------------------------------------------------------
def Config(algorithm=ALGO_AUTO, chunk_combos=0, candidate_cap=0, stream=None, sample_threshold=True,
           mfma_int8=False, variant=0):
    cfg = _lib.Config()
    lib().fiu_config_init(C.byref(cfg))
    cfg.algorithm, cfg.chunk_combos, cfg.candidate_cap = algorithm, chunk_combos, candidate_cap
    cfg.flags = (0 if sample_threshold else _lib.FLAG_NO_THRESHOLD_SAMPLE) | (_lib.FLAG_MFMA_INT8 if mfma_int8 else 0)
    cfg.stream = stream
    cfg.variant = variant
    return cfg


class GpuSearch:
    """Search on one MI355X (drop-in for ThreadedSearch, ThreadedSearch.h:28)."""

    def __init__(self, algorithm=ALGO_AUTO, chunk_combos=0, candidate_cap=0, sample_threshold=True,
                 mfma_int8=False, variant=0):
        self.config = Config(algorithm, chunk_combos, candidate_cap, sample_threshold=sample_threshold,
                             mfma_int8=mfma_int8, variant=variant)
        self.last_stats = None

    def run_arrays(self, dataset: Dataset, order: int, distribution: Distribution, outputs: int):
        if distribution.k != order - 1:
            raise ValueError("distribution.k must be order - 1")
        combos = np.zeros(max(outputs, 1) * order, np.int32)
        mi = np.zeros(max(outputs, 1), np.float32)
        cnt = C.c_uint32()
        stats = _lib.SearchStats()
        check(lib().fiu_search(dataset.handle, order, distribution.step, distribution.offset,
                               outputs, C.byref(self.config), combos.ctypes.data, mi.ctypes.data,
                               C.byref(cnt), C.byref(stats)))
        self.last_stats = stats
        n = cnt.value
        return combos[:n * order].reshape(n, order), mi[:n]

    def run(self, dataset: Dataset, order: int, distribution: Distribution, outputs: int):
        """``Search::run``: results sorted by MI descending (ties by combination)."""
        combos, mi = self.run_arrays(dataset, order, distribution, outputs)
        return [Result([int(x) for x in c], float(v)) for c, v in zip(combos, mi)]


# ---------------------------------------------------------------------------
# Batched searches and the scan server
# ---------------------------------------------------------------------------
def batch_config(max_active=0, memory_budget=0, step_ms=0.0, candidate_cap=0):
    cfg = _lib.BatchConfig()
    lib().fiu_batch_config_init(C.byref(cfg))
    if max_active:
        cfg.max_active = max_active
    if memory_budget:
        cfg.memory_budget = memory_budget
    if step_ms:
        cfg.step_ms = step_ms
    if candidate_cap:
        cfg.candidate_cap = candidate_cap
    return cfg


def search_batch(scans, config=None):
    """Run many searches together (``fiu_search_batch``).

    ``scans``: sequence of (dataset, order, distribution, outputs).  Returns
    (list of (combos, mi) per search, batch stats)."""
    n = len(scans)
    arr = (_lib.Scan * max(n, 1))()
    bufs = []
    for i, (ds, order, dist, outputs) in enumerate(scans):
        if dist.k != order - 1:
            raise ValueError("distribution.k must be order - 1")
        combos = np.zeros(max(outputs, 1) * order, np.int32)
        mi = np.zeros(max(outputs, 1), np.float32)
        bufs.append((combos, mi, order))
        s = arr[i]
        s.dataset, s.order, s.dist_step, s.dist_offset, s.outputs = (
            ds.handle.value, order, dist.step, dist.offset, outputs)
        s.out_combos, s.out_mi = combos.ctypes.data, mi.ctypes.data
    stats = _lib.BatchStats()
    cfg = config if config is not None else batch_config()
    check(lib().fiu_search_batch(arr, n, C.byref(cfg), C.byref(stats)))
    out = []
    for i, (combos, mi, order) in enumerate(bufs):
        c = arr[i].out_count
        out.append((combos[:c * order].reshape(c, order), mi[:c]))
    return out, stats


class Server:
    """Scan server (``fiu_server_*``): continuous batching of scan requests."""

    def __init__(self, config=None):
        cfg = config if config is not None else batch_config()
        h = C.c_void_p()
        check(lib().fiu_server_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self._pending = {}

    def submit(self, order, outputs, dataset=None, words=None, snps=0, cases_words=0, ctrls_words=0,
               cases=0, ctrls=0, distribution=None):
        """Submit a request on a resident ``dataset`` or on host ``words``
        (reference layout; copied before this returns).  Returns its id."""
        r = _lib.Request()
        r.struct_size = C.sizeof(_lib.Request)
        r.order, r.outputs = order, outputs
        dist = distribution or Distribution(0, order - 1)
        r.dist_step, r.dist_offset = dist.step, dist.offset
        if dataset is not None:
            r.dataset = dataset.handle.value
        else:
            w = np.ascontiguousarray(words, dtype=np.uint64)
            if w.size != snps * 3 * (cases_words + ctrls_words):
                raise ValueError("words has the wrong length for the given shape")
            r.words = w.ctypes.data
            r.snps, r.cases_words, r.ctrls_words, r.cases, r.ctrls = snps, cases_words, ctrls_words, cases, ctrls
        rid = C.c_uint64()
        check(lib().fiu_server_submit(self._h, C.byref(r), C.byref(rid)))
        self._pending[rid.value] = (order, outputs)
        return rid.value

    def wait(self, rid, timeout_ms=-1.0):
        """(combos, mi, stats) of request ``rid``; None when still running
        after ``timeout_ms`` (>= 0)."""
        order, outputs = self._pending[rid]
        combos = np.zeros(max(outputs, 1) * order, np.int32)
        mi = np.zeros(max(outputs, 1), np.float32)
        cnt = C.c_uint32()
        st = _lib.RequestStats()
        s = lib().fiu_server_wait(self._h, rid, timeout_ms, combos.ctypes.data, mi.ctypes.data, C.byref(cnt),
                                  C.byref(st))
        if s == _lib.FIU_ERR_PENDING:
            return None
        del self._pending[rid]
        check(s)
        n = cnt.value
        return combos[:n * order].reshape(n, order), mi[:n], st

    def info(self):
        i = _lib.ServerInfo()
        check(lib().fiu_server_query(self._h, C.byref(i)))
        return i

    def close(self):
        if self._h and self._h.value:
            check(lib().fiu_server_destroy(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Batch operators
# ---------------------------------------------------------------------------
def combine_batch(dataset: Dataset, combos) -> np.ndarray:
    """Order-m genotype tables of each combination (n x 3^m rows x words)."""
    combos = np.ascontiguousarray(combos, dtype=np.int32)
    n, m = combos.shape
    per = 3 ** m * (dataset.cases_words + dataset.ctrls_words)
    out = np.zeros(max(n * per, 1), np.uint64)
    check(lib().fiu_combine_batch(dataset.handle, combos.ctypes.data, m, n, out.ctypes.data))
    return out[:n * per].reshape(n, per)


def contingency_batch(dataset: Dataset, combos, with_mi=True):
    """(counts n x 2 x ct_size, mi n) of each combination."""
    combos = np.ascontiguousarray(combos, dtype=np.int32)
    n, k = combos.shape
    cts = ct_size(k)
    counts = np.zeros(max(n * 2 * cts, 1), np.uint32)
    mi = np.zeros(max(n, 1), np.float32)
    check(lib().fiu_contingency_batch(dataset.handle, combos.ctypes.data, k, n, counts.ctypes.data,
                                      mi.ctypes.data if with_mi else None))
    return counts[:n * 2 * cts].reshape(n, 2, cts), mi[:n]


def mi_batch(tables, order, cases, ctrls) -> np.ndarray:
    """MI of caller-given tables (n x 2 x ct_size(order))."""
    tables = np.ascontiguousarray(tables, dtype=np.uint32)
    n = tables.shape[0]
    out = np.zeros(max(n, 1), np.float32)
    check(lib().fiu_mi_batch(tables.ctypes.data, n, order, cases, ctrls, out.ctypes.data))
    return out[:n]

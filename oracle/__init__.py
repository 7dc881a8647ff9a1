"""ctypes bindings of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this package, and only as the checker (or the timed CPU
baseline).  The product package ``fiuncho_amd`` never imports it.

``liboracle.so`` is our plain-C restatement of the reference search path
(oracle.c cites the reference file:line of every function).  ``_ref/`` holds
the reference itself, compiled from its own sources by ``oracle/Makefile``
when they are present; :func:`ref` returns ``None`` when it is not built.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = {}

u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
i8p = np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")


class _DS(C.Structure):
    _fields_ = [("snps", C.c_size_t), ("cases", C.c_size_t), ("ctrls", C.c_size_t),
                ("cases_words", C.c_size_t), ("ctrls_words", C.c_size_t),
                ("words", C.POINTER(C.c_uint64))]


def build():
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_words_for.restype = C.c_size_t
        L.orc_words_for.argtypes = [C.c_size_t, C.c_int]
        L.orc_ct_size.restype = C.c_size_t
        L.orc_ct_size.argtypes = [C.c_int, C.c_int]
        L.orc_read_raw.argtypes = [C.c_char_p, C.c_int, C.POINTER(_DS), C.c_char_p, C.c_size_t]
        L.orc_read_tped.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.POINTER(_DS), C.c_char_p, C.c_size_t]
        L.orc_pack.argtypes = [i8p, u8p, C.c_size_t, C.c_size_t, C.c_int, C.POINTER(_DS)]
        L.orc_dataset_free.argtypes = [C.POINTER(_DS)]
        L.orc_combine.argtypes = [u64p, C.c_size_t, u64p, C.c_size_t, u64p]
        L.orc_popcnt.argtypes = [u64p, C.c_size_t, u64p, C.c_size_t, u32p, C.c_size_t]
        L.orc_counts.argtypes = [C.POINTER(_DS), i32p, C.c_int, u32p, u32p]
        L.orc_counts_bruteforce.argtypes = [i8p, u8p, C.c_size_t, i32p, C.c_int, u32p, u32p]
        for f in ("orc_mi_f32_ref512", "orc_mi_f32_base"):
            getattr(L, f).restype = C.c_float
            getattr(L, f).argtypes = [u32p, u32p, C.c_size_t, C.c_uint, C.c_uint]
        L.orc_mi_f64.restype = C.c_double
        L.orc_mi_f64.argtypes = [u32p, u32p, C.c_size_t, C.c_uint, C.c_uint]
        L.orc_dist_prefixes.restype = C.c_size_t
        L.orc_dist_prefixes.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_size_t]
        L.orc_search.restype = C.c_long
        L.orc_search.argtypes = [C.POINTER(_DS), C.c_int, C.c_int, C.c_int, C.c_uint, C.c_int,
                                 C.c_int, i32p, f32p, C.POINTER(C.c_uint64)]
        L.orc_has_avx512.restype = C.c_int
        _LIB = L
    return _LIB


class Dataset:
    """A dataset in the reference host layout, owned by the oracle."""

    def __init__(self, ds: _DS):
        self._ds = ds
        self.snps, self.cases, self.ctrls = ds.snps, ds.cases, ds.ctrls
        self.cases_words, self.ctrls_words = ds.cases_words, ds.ctrls_words
        n = self.snps * 3 * (self.cases_words + self.ctrls_words)
        self.words = np.ctypeslib.as_array(ds.words, shape=(n,)).copy() if n else np.zeros(0, np.uint64)

    def __del__(self):
        try:
            lib().orc_dataset_free(C.byref(self._ds))
        except Exception:
            pass

    def snp(self, s):
        per = 3 * (self.cases_words + self.ctrls_words)
        t = self.words[s * per:(s + 1) * per]
        return t[:3 * self.cases_words], t[3 * self.cases_words:]


class WordsDataset:
    """Tables already in the reference host layout (a caller's array, kept
    alive here), for counts() and search()."""

    def __init__(self, words, snps, cases, ctrls, cases_words, ctrls_words):
        self.words = np.ascontiguousarray(words, dtype=np.uint64)
        assert self.words.size == snps * 3 * (cases_words + ctrls_words)
        self.snps, self.cases, self.ctrls = snps, cases, ctrls
        self.cases_words, self.ctrls_words = cases_words, ctrls_words


def from_words(words, snps, cases, ctrls, cases_words, ctrls_words):
    return WordsDataset(words, snps, cases, ctrls, cases_words, ctrls_words)


def _err():
    return C.create_string_buffer(512)


def read_raw(path, align=64):
    ds, e = _DS(), _err()
    if lib().orc_read_raw(path.encode(), align, C.byref(ds), e, 512) != 0:
        raise RuntimeError(e.value.decode())
    return Dataset(ds)


def read_tped(tped, tfam, align=64):
    ds, e = _DS(), _err()
    if lib().orc_read_tped(tped.encode(), tfam.encode(), align, C.byref(ds), e, 512) != 0:
        raise RuntimeError(e.value.decode())
    return Dataset(ds)


def pack(geno, is_ctrl, align=64):
    geno = np.ascontiguousarray(geno, dtype=np.int8)
    is_ctrl = np.ascontiguousarray(is_ctrl, dtype=np.uint8)
    ds = _DS()
    if lib().orc_pack(geno, is_ctrl, geno.shape[0], geno.shape[1], align, C.byref(ds)) != 0:
        raise MemoryError
    return Dataset(ds)


def _as_ds(d: Dataset):
    ds = _DS(d.snps, d.cases, d.ctrls, d.cases_words, d.ctrls_words,
             d.words.ctypes.data_as(C.POINTER(C.c_uint64)))
    return ds


def counts(d: Dataset, combo):
    k = len(combo)
    cells = 3 ** k
    cc, ct = np.zeros(cells, np.uint32), np.zeros(cells, np.uint32)
    lib().orc_counts(C.byref(_as_ds(d)), np.asarray(combo, np.int32), k, cc, ct)
    return cc, ct


def counts_bruteforce(geno, is_ctrl, combo):
    geno = np.ascontiguousarray(geno, dtype=np.int8)
    is_ctrl = np.ascontiguousarray(is_ctrl, dtype=np.uint8)
    k = len(combo)
    cc, ct = np.zeros(3 ** k, np.uint32), np.zeros(3 ** k, np.uint32)
    lib().orc_counts_bruteforce(geno, is_ctrl, geno.shape[1], np.asarray(combo, np.int32), k, cc, ct)
    return cc, ct


def _padded(cc, ct, k):
    size = lib().orc_ct_size(k, 64)
    a, b = np.zeros(size, np.uint32), np.zeros(size, np.uint32)
    a[:len(cc)], b[:len(ct)] = cc, ct
    return a, b, size


def mi_f32_ref512(cc, ct, k, cases, ctrls):
    a, b, size = _padded(cc, ct, k)
    return float(lib().orc_mi_f32_ref512(a, b, size, cases, ctrls))


def mi_f32_base(cc, ct, k, cases, ctrls):
    a, b, size = _padded(cc, ct, k)
    return float(lib().orc_mi_f32_base(a, b, size, cases, ctrls))


def mi_f64(cc, ct, k, cases, ctrls):
    a, b, size = _padded(cc, ct, k)
    return float(lib().orc_mi_f64(a, b, size, cases, ctrls))


def dist_prefixes(n, k1, step, offset):
    cnt = lib().orc_dist_prefixes(n, k1, step, offset, None, 0)
    out = np.zeros(max(cnt, 1) * k1, np.int32)
    lib().orc_dist_prefixes(n, k1, step, offset, out.ctypes.data, cnt)
    return out[:cnt * k1].reshape(cnt, k1)


def search(d: Dataset, order, outputs, step=1, offset=0, threads=1, mi_mode=2):
    """Top-``outputs`` (combos, mi) ordered by (mi desc, combination asc).

    mi_mode 0: FP32 in the reference AVX-512 order, 1: FP32 scalar order,
    2: FP64 rounded to float."""
    combos = np.zeros(max(outputs, 1) * order, np.int32)
    mi = np.zeros(max(outputs, 1), np.float32)
    ev = C.c_uint64(0)
    n = lib().orc_search(C.byref(_as_ds(d)), order, step, offset, outputs, threads, mi_mode,
                         combos, mi, C.byref(ev))
    if n < 0:
        raise RuntimeError("orc_search failed")
    return combos[:n * order].reshape(n, order), mi[:n], ev.value


def has_avx512():
    return bool(lib().orc_has_avx512())


# ---------------------------------------------------------------------------
# The reference compiled from source (oracle/_ref), when present
# ---------------------------------------------------------------------------
class Ref:
    def __init__(self, path):
        L = C.CDLL(path, mode=os.RTLD_LOCAL if hasattr(os, "RTLD_LOCAL") else 0)
        L.ref_last_error.restype = C.c_char_p
        L.ref_load.restype = C.c_void_p
        L.ref_load.argtypes = [C.c_char_p, C.c_char_p]
        L.ref_free.argtypes = [C.c_void_p]
        L.ref_info.argtypes = [C.c_void_p] + [C.POINTER(C.c_size_t)] * 5
        L.ref_words.argtypes = [C.c_void_p, u64p]
        L.ref_combination.restype = C.c_size_t
        L.ref_combination.argtypes = [C.c_void_p, i32p, C.c_int, u32p, C.POINTER(C.c_float)]
        L.ref_combine_pair.argtypes = [C.c_void_p, C.c_int, C.c_int, u64p]
        L.ref_mi.restype = C.c_float
        L.ref_mi.argtypes = [u32p, u32p, C.c_int, C.c_uint, C.c_uint]
        for f in ("ref_result_str", "ref_result_serialize"):
            getattr(L, f).restype = C.c_size_t
            getattr(L, f).argtypes = [i32p, C.c_int, C.c_float, C.c_char_p, C.c_size_t]
        L.ref_search.restype = C.c_long
        L.ref_search.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_uint, C.c_uint,
                                 i32p, f32p, C.POINTER(C.c_double)]
        self.L = L

    def load(self, f1, f2=None):
        h = self.L.ref_load(f1.encode(), (f2 or "").encode())
        if not h:
            raise RuntimeError(self.L.ref_last_error().decode())
        return RefDataset(self, h)

    def result_str(self, combo, val):
        """Result<int, float>::str of one result (bytes)."""
        c = np.ascontiguousarray(combo, np.int32)
        n = self.L.ref_result_str(c, len(c), val, None, 0)
        buf = C.create_string_buffer(n + 1)
        self.L.ref_result_str(c, len(c), val, buf, n)
        return buf.raw[:n]

    def result_serialize(self, combo, val):
        """Result<int, float>::serialize of one result (bytes)."""
        c = np.ascontiguousarray(combo, np.int32)
        n = self.L.ref_result_serialize(c, len(c), val, None, 0)
        buf = C.create_string_buffer(n)
        self.L.ref_result_serialize(c, len(c), val, buf, n)
        return buf.raw[:n]

    def mi(self, cc, ct, k, cases, ctrls):
        return float(self.L.ref_mi(np.ascontiguousarray(cc, np.uint32),
                                   np.ascontiguousarray(ct, np.uint32), k, cases, ctrls))


class RefDataset:
    def __init__(self, ref, h):
        self.ref, self.h = ref, h
        v = [C.c_size_t() for _ in range(5)]
        ref.L.ref_info(h, *[C.byref(x) for x in v])
        self.snps, self.cases, self.ctrls, self.cases_words, self.ctrls_words = [x.value for x in v]

    def __del__(self):
        try:
            self.ref.L.ref_free(self.h)
        except Exception:
            pass

    def words(self):
        out = np.zeros(self.snps * 3 * (self.cases_words + self.ctrls_words), np.uint64)
        self.ref.L.ref_words(self.h, out)
        return out

    def combination(self, combo):
        k = len(combo)
        size = lib().orc_ct_size(k, 64)
        ct = np.zeros(2 * size, np.uint32)
        mi = C.c_float()
        self.ref.L.ref_combination(self.h, np.asarray(combo, np.int32), k, ct, C.byref(mi))
        return ct[:3 ** k], ct[size:size + 3 ** k], float(mi.value)

    def combine_pair(self, a, b):
        out = np.zeros(9 * (self.cases_words + self.ctrls_words), np.uint64)
        self.ref.L.ref_combine_pair(self.h, a, b, out)
        return out

    def search(self, order, outputs, step=1, offset=0, threads=1):
        combos = np.zeros(max(outputs, 1) * order, np.int32)
        mi = np.zeros(max(outputs, 1), np.float32)
        sec = C.c_double()
        n = self.ref.L.ref_search(self.h, order, step, offset, outputs, threads, combos, mi,
                                  C.byref(sec))
        if n < 0:
            raise RuntimeError(self.ref.L.ref_last_error().decode())
        return combos[:n * order].reshape(n, order), mi[:n], sec.value


def ref(kind="avx512"):
    """The reference built from source, or None when oracle/_ref lacks it."""
    name = {"avx512": "libfiuncho_ref.so", "base": "libfiuncho_ref_base.so"}[kind]
    if kind not in _REF:
        path = os.path.join(HERE, "_ref", name)
        _REF[kind] = Ref(path) if os.path.exists(path) else None
    return _REF[kind]

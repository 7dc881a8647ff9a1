"""The C ABI library: it loads, exports every entry point include/fiuncho_amd.h
declares, and its host-only parts (generator, sizes) behave, with no GPU."""
import ctypes as C
import os
import re

import numpy as np

import fiuncho_amd as fa
from fiuncho_amd import _lib

from .conftest import ROOT

HEADER = os.path.join(ROOT, "include", "fiuncho_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fiu_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    decl = declared_functions()
    assert decl, "no declarations parsed"
    for name in decl:
        assert hasattr(L, name), name
    assert sorted(_lib.EXPORTS) == decl


def test_host_only_entry_points():
    assert "gfx950" in fa.version()
    assert [fa.ct_size(k) for k in (1, 2, 3, 4, 5)] == [16, 16, 32, 96, 256]
    cfg = fa.Config()
    assert cfg.struct_size == C.sizeof(_lib.Config) and cfg.algorithm == fa.ALGO_AUTO


def test_synth_deterministic_and_packed_like_the_oracle():
    import oracle as O
    spec = fa.synth_spec(37, 130, 71, seed=7, planted=(2, 30))
    w1, g1, cw, tw = fa.synth_generate(spec, words=True, geno=True)
    spec.threads = 1
    w2, g2, _, _ = fa.synth_generate(spec, words=True, geno=True)
    assert np.array_equal(w1, w2) and np.array_equal(g1, g2)
    assert (cw, tw) == (8, 8)
    is_ctrl = np.r_[np.zeros(130, np.uint8), np.ones(71, np.uint8)]
    d = O.pack(g1, is_ctrl)
    assert np.array_equal(d.words, w1)


def test_synth_hwe_frequencies():
    spec = fa.synth_spec(200, 4000, 4000, seed=3)
    _, g, _, _ = fa.synth_generate(spec, words=False, geno=True)
    p = g.mean(axis=1) / 2.0                   # allele frequency estimate
    assert (p > 0.03).all() and (p < 0.53).all()
    het = (g == 1).mean(axis=1)
    assert np.abs(het - 2 * p * (1 - p)).max() < 0.04

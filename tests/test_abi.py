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


def test_product_loader_matches_reference_tables(tmp_path):
    import hashlib
    import json
    from .conftest import DATA, GOLDEN
    with open(os.path.join(GOLDEN, "testdata_search.json")) as f:
        sha = json.load(f)["sha256"]
    for files in ([os.path.join(DATA, "test.tped"), os.path.join(DATA, "test.tfam")],
                  [os.path.join(DATA, "test.tfam"), os.path.join(DATA, "test.tped")],
                  [os.path.join(DATA, "test.raw")], [os.path.join(DATA, "test.gametes.raw")]):
        words, snps, cases, ctrls, cw, tw = fa.read_files(files)
        assert (snps, cases, ctrls, cw, tw) == (10, 9000, 9800, 144, 160)
        assert hashlib.sha256(words.tobytes()).hexdigest() == sha["tped"]
    q = np.load(os.path.join(GOLDEN, "loader_quirk.npz"))
    for key, name in (("raw_text", "q.raw"), ("tped_text", "q.tped"), ("tfam_text", "q.tfam")):
        (tmp_path / name).write_bytes(q[key].tobytes())
    assert np.array_equal(fa.read_files([str(tmp_path / "q.raw")])[0], q["raw_words"])
    assert np.array_equal(fa.read_files([str(tmp_path / "q.tped"), str(tmp_path / "q.tfam")])[0], q["tped_words"])


def test_product_loader_errors(tmp_path):
    import pytest
    from .conftest import DATA
    bad = tmp_path / "x.txt"
    bad.write_text("1 2 3\n")
    for files in ([str(bad)], [os.path.join(DATA, "test.tped")], [os.path.join(DATA, "test.tped")] * 2,
                  [str(tmp_path / "missing.raw")], []):
        with pytest.raises(fa.FiuError):
            fa.read_files(files)
    tfam = tmp_path / "b.tfam"
    tfam.write_text("f1 i1 0 0 1 3\n")          # phenotype must be 1 or 2
    tped = tmp_path / "b.tped"
    tped.write_text("1 s 0 1 A A\n")
    with pytest.raises(fa.FiuError):
        fa.read_files([str(tped), str(tfam)])
    tfam.write_text("f1 i1 0 0 1 2\n")
    tped.write_text("1 s 0 1 A X\n")            # invalid nucleotide
    with pytest.raises(fa.FiuError):
        fa.read_files([str(tped), str(tfam)])
    tped.write_text("1 s 0 1 A A C C\n")        # 2 samples' alleles for 1 sample
    with pytest.raises(fa.FiuError):
        fa.read_files([str(tped), str(tfam)])


def test_tuning_table_roundtrip_and_validation():
    from fiuncho_amd import tune
    E = tune.Entry
    good = [E(2, 0, 1000, 0, 4096, fa.ALGO_MARGINAL, 8, 1.5), E(3, 100, 200, 10, 20, fa.ALGO_MFMA, 3),
            E(4, 0, 50, 0, 100, fa.ALGO_DIRECT, 0), E(5, 0, 50, 0, 100, fa.ALGO_MARGINAL, 0),
            E(7, 0, 50, 0, 100, fa.ALGO_DIRECT, 0), E(6, 0, 50, 0, 100, fa.ALGO_DIRECT, 0),
            E(4, 50, 500, 0, 100, fa.ALGO_MARGINAL, 0, 0.0, 6), E(2, 50, 500, 0, 100, fa.ALGO_MFMA, 2, 0.0, 2)]
    try:
        tune.set_table(good)
        got = tune.get_table()
        assert [(e.order, e.snps_lo, e.snps_hi, e.samples_lo, e.samples_hi, e.algorithm, e.variant) for e in got] == \
               [(e.order, e.snps_lo, e.snps_hi, e.samples_lo, e.samples_hi, e.algorithm, e.variant) for e in good]
        assert abs(got[0].ms - 1.5) < 1e-6
        assert [e.batch_unit for e in got] == [e.batch_unit for e in good]
        bad = [E(3, 0, 10, 0, 10, fa.ALGO_MARGINAL, 8),     # 8 prefixes per wave: order 2 only
               E(2, 0, 10, 0, 10, fa.ALGO_MFMA, 5),         # ring depth: order 3 only
               E(3, 0, 10, 0, 10, fa.ALGO_MFMA, 7),         # no such tile / staging variant
               E(4, 0, 10, 0, 10, fa.ALGO_MFMA, 0),         # matrix cores: orders 2-3
               E(13, 0, 10, 0, 10, fa.ALGO_DIRECT, 0),      # orders 2-12 only
               E(7, 0, 10, 0, 10, fa.ALGO_MARGINAL, 0),     # marginal: orders 2-6
               E(10, 0, 10, 0, 10, fa.ALGO_MARGINAL, 0),    # marginal above order 6
               E(4, 0, 10, 0, 10, fa.ALGO_MARGINAL, 4),     # prefixes per wave: orders 2-3
               E(2, 10, 10, 0, 10, fa.ALGO_DIRECT, 0),      # empty SNP range
               E(2, 0, 10, 0, 10, fa.ALGO_AUTO, 0),         # AUTO is not a choice
               E(1, 0, 10, 0, 10, fa.ALGO_DIRECT, 0),
               E(4, 0, 10, 0, 10, fa.ALGO_MARGINAL, 0, 0.0, 7),     # batch unit 0..6
               E(3, 0, 10, 0, 10, fa.ALGO_MFMA, 0, 0.0, 3)]         # batch tile 0..2
        for b in bad:
            try:
                tune.set_table([b])
            except fa.FiuError as e:
                assert e.status == _lib.FIU_ERR_INVALID
            else:
                raise AssertionError(f"accepted {b}")
        assert len(tune.get_table()) == len(good)   # a refused table leaves the old one
    finally:
        tune.clear()
    assert tune.get_table() == []
    # the autotuner's candidates are all valid table entries
    for k in (2, 3):
        for a, v in tune.candidates(k):
            tune.set_table([E(k, 0, 10, 0, 10, a, v)])
    tune.clear()


def test_genotype_ingest_validation_before_any_device_call():
    import pytest
    g = np.zeros((3, 4), np.int8)
    for geno, is_ctrl, flags in ((g, np.array([0, 1, 0, 1], np.uint8), 2),        # unknown flag
                                 (np.zeros((3, 0), np.int8), np.zeros(0, np.uint8), 0),   # no samples
                                 (g, np.zeros(4, np.uint8), 0),                   # no controls
                                 (g, np.ones(4, np.uint8), 0)):                   # no cases
        h = C.c_void_p()
        with pytest.raises(fa.FiuError) as e:
            _lib.check(_lib.lib().fiu_dataset_from_genotypes(geno.ctypes.data, geno.shape[0], geno.shape[1],
                                                             is_ctrl.ctypes.data, flags, C.byref(h)))
        assert e.value.status == _lib.FIU_ERR_INVALID
    with pytest.raises(ValueError):
        fa.Dataset.from_genotypes(g, np.zeros(3, np.uint8))


def test_parallel_parse_reports_the_first_bad_line(tmp_path):
    # files large enough to be parsed on several threads: the error names the
    # first bad line of the file, as a line-by-line reader would
    import pytest
    rng = np.random.default_rng(4)
    g = rng.integers(0, 3, size=(3000, 5))
    rows = [" ".join(["f", "i", "0", "0", "1", str(1 + (j % 2))] + [str(v) for v in g[j]]) for j in range(3000)]
    rows[2700] = rows[2700].replace(" 1 ", " NA ", 1) if " 1 " in rows[2700] else rows[2700] + " x"
    rows[1700] = rows[1700].rsplit(" ", 1)[0] + " NA"
    p = tmp_path / "big.raw"
    p.write_text("FID IID PAT MAT SEX PHENOTYPE " + " ".join(f"S{i}_A" for i in range(5)) + "\n" + "\n".join(rows) + "\n")
    with pytest.raises(fa.FiuError, match="at line 1702$"):
        fa.read_files([str(p)])
    rows[1700] = rows[1700].rsplit(" ", 1)[0] + " 2"
    rows[2700] = " ".join(["f", "i", "0", "0", "1", "1"] + ["0"] * 5)
    p.write_text("FID IID PAT MAT SEX PHENOTYPE " + " ".join(f"S{i}_A" for i in range(5)) + "\n" + "\n".join(rows) + "\n")
    words, snps, cases, ctrls, _, _ = fa.read_files([str(p)])
    assert (snps, cases, ctrls) == (5, 1500, 1500)
    tfam = tmp_path / "b.tfam"
    tfam.write_text("".join(f"f{k} i{k} 0 0 1 {1 + k % 2}\n" for k in range(4)))
    lines = [f"1 rs{s} 0 {s} A C A C A A C C" for s in range(1200)]
    lines[1000] = "1 rs 0 1 A C A C A A C"       # odd number of nucleotides
    lines[600] = "1 rs 0 1 A C A C A A C N"      # invalid nucleotide
    tped = tmp_path / "b.tped"
    tped.write_text("\n".join(lines) + "\n")
    with pytest.raises(fa.FiuError, match=r"b\.tped:601: invalid nucleotide value 'N'"):
        fa.read_files([str(tped), str(tfam)])

"""GPU: device-side packing of genotype matrices (fiu_dataset_from_genotypes,
and fiu_dataset_read, which parses on the host and packs on the device).

Bar: the device-packed words equal, bit for bit, the oracle's packing of the
same genotypes (RawFile::fill layout, RAWFile.hpp:209-290), the words the
reference reader builds from the reference's own test files (sha256 of the
tables, tests/dataset.cpp sizes 144/160 words) and from the re-flush fixture,
and a search over the packed dataset returns the oracle's top-N."""
import hashlib
import json
import os

import numpy as np
import pytest

import fiuncho_amd as fa
import oracle as O

from .conftest import DATA, GOLDEN

pytestmark = pytest.mark.gpu


def genotypes(snps, samples, seed, uncalled=0.0, ctrl_frac=0.5, sorted_classes=False):
    rng = np.random.default_rng(seed)
    geno = rng.integers(0, 3, size=(snps, samples)).astype(np.int8)
    if uncalled:
        hole = rng.random((snps, samples)) < uncalled
        geno[hole] = rng.choice(np.array([3, -1, 7, 100], np.int8), size=int(hole.sum()))
    is_ctrl = (rng.random(samples) < ctrl_frac).astype(np.uint8)
    if sorted_classes:
        is_ctrl = np.sort(is_ctrl)
    return geno, is_ctrl


@pytest.mark.parametrize("snps,samples,seed,uncalled,frac,srt", [
    (7, 200, 1, 0.0, 0.5, False),       # interleaved classes, one partial word each
    (33, 1000, 2, 0.05, 0.3, False),    # uncalled values, uneven classes
    (5, 128, 3, 0.0, 0.5, True),        # cases then controls, whole words
    (64, 1025, 4, 0.01, 0.5, False),    # class sizes around 512 (padding words)
    (3, 2, 5, 0.0, 0.5, True),          # one sample per class
    (300, 4096, 6, 0.0, 0.5, False),    # many word slots
    (3, 50000, 7, 0.01, 0.5, False),    # rows too long to stage in LDS, interleaved classes
    (4, 60001, 8, 0.0, 0.4, True),      # rows too long to stage in LDS, class runs
    (9, 16001, 10, 0.0, 0.5, True),     # staged rows not starting at 16-byte boundaries
])
def test_device_pack_matches_the_oracle(snps, samples, seed, uncalled, frac, srt):
    print(f"case: pack {snps}x{samples} seed {seed}", flush=True)
    geno, is_ctrl = genotypes(snps, samples, seed, uncalled, frac, srt)
    if is_ctrl.all() or not is_ctrl.any():
        is_ctrl[0] ^= 1
    od = O.pack(geno, is_ctrl)
    d = fa.Dataset.from_genotypes(geno, is_ctrl)
    assert (d.snps, d.cases, d.ctrls, d.cases_words, d.ctrls_words) == \
           (snps, od.cases, od.ctrls, od.cases_words, od.ctrls_words)
    assert np.array_equal(d.words(), od.words)
    # the device image is built from the packed words: a search agrees
    if snps >= 3:
        oc, om, _ = O.search(od, 2, 20, threads=4, mi_mode=2)
        c, m = fa.GpuSearch().run_arrays(d, 2, fa.Distribution(snps, 1), 20)
        assert c.tolist() == oc.tolist() and np.array_equal(m, om)
    d.close()


def test_read_packs_on_the_device_like_the_reference_reader(tmp_path):
    with open(os.path.join(GOLDEN, "testdata_search.json")) as f:
        sha = json.load(f)["sha256"]["tped"]
    for files in ([os.path.join(DATA, "test.tped"), os.path.join(DATA, "test.tfam")],
                  [os.path.join(DATA, "test.raw")], [os.path.join(DATA, "test.gametes.raw")]):
        print(f"case: read {os.path.basename(files[0])}", flush=True)
        d = fa.Dataset.read(files)
        assert (d.snps, d.cases, d.ctrls, d.cases_words, d.ctrls_words) == (10, 9000, 9800, 144, 160)
        assert hashlib.sha256(d.words().tobytes()).hexdigest() == sha
        d.close()
    # the readers' second flush of an emptied buffer (a class reaching 64k
    # samples right before a sample of the other class)
    q = np.load(os.path.join(GOLDEN, "loader_quirk.npz"))
    for key, name in (("raw_text", "q.raw"), ("tped_text", "q.tped"), ("tfam_text", "q.tfam")):
        (tmp_path / name).write_bytes(q[key].tobytes())
    assert np.array_equal(fa.Dataset.read([str(tmp_path / "q.raw")]).words(), q["raw_words"])
    assert np.array_equal(fa.Dataset.read([str(tmp_path / "q.tfam"), str(tmp_path / "q.tped")]).words(),
                          q["tped_words"])


def test_reader_flush_flag_zeroes_exactly_the_reflushed_words():
    # classes alternate in runs so that several class-64k boundaries are
    # followed by the other class; the flag zeroes those words (all 3 rows)
    print("case: reader flush words", flush=True)
    rng = np.random.default_rng(9)
    is_ctrl = np.r_[np.zeros(64, np.uint8), np.ones(70, np.uint8), np.zeros(64, np.uint8),
                    np.ones(58, np.uint8), np.zeros(10, np.uint8)]
    geno = rng.integers(0, 3, size=(6, is_ctrl.size)).astype(np.int8)
    plain = fa.Dataset.from_genotypes(geno, is_ctrl).words()
    flushed = fa.Dataset.from_genotypes(geno, is_ctrl, reader_flush=True).words()
    od = O.pack(geno, is_ctrl)
    assert np.array_equal(plain, od.words)
    cw, tw = od.cases_words, od.ctrls_words
    expect = od.words.copy().reshape(6, -1)
    # case word 0 ends at sample 63 (next is a control), control word 1 ends
    # at sample 64+70+64+57 = 255 (next is a case); case word 1 ends at
    # sample 64+70+63 (next is a control)
    for cls, w in ((0, 0), (1, 1), (0, 1)):
        W, off = (tw, 3 * cw) if cls else (cw, 0)
        for z in range(3):
            expect[:, off + z * W + w] = 0
    assert np.array_equal(flushed, expect.ravel())


def test_no_snps():
    geno = np.zeros((0, 10), np.int8)
    d = fa.Dataset.from_genotypes(geno, np.r_[np.zeros(5, np.uint8), np.ones(5, np.uint8)])
    assert (d.snps, d.cases, d.ctrls) == (0, 5, 5)
    assert d.words().size == 0
    d.close()

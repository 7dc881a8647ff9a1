"""The oracle pinned against the reference's own fixtures, known answers, the
golden vectors written by the reference compiled from source, and closed forms.

Each test names the reference test or file:line it restates.
"""
import itertools
import json
import math
import os

import numpy as np
import pytest

import oracle as O

from .conftest import DATA, GOLDEN

TPED, TFAM = os.path.join(DATA, "test.tped"), os.path.join(DATA, "test.tfam")
RAW, GAMETES = os.path.join(DATA, "test.raw"), os.path.join(DATA, "test.gametes.raw")

# tests/utils.h:49-84: MPI3SNP top-100 triples on tests/data
MPI3SNP = [
    ((4, 8, 9), 0.198944), ((6, 8, 9), 0.198866), ((3, 8, 9), 0.198622), ((7, 8, 9), 0.198564),
    ((5, 8, 9), 0.198526), ((0, 8, 9), 0.198519), ((1, 8, 9), 0.198300), ((2, 8, 9), 0.198216),
    ((5, 6, 7), 0.001916), ((0, 5, 7), 0.001585), ((0, 4, 5), 0.001559), ((1, 6, 7), 0.001548),
    ((1, 2, 7), 0.001522), ((3, 5, 8), 0.001495), ((4, 6, 7), 0.001482), ((6, 7, 8), 0.001478),
    ((1, 4, 5), 0.001474), ((0, 1, 9), 0.001451), ((0, 4, 9), 0.001429), ((3, 5, 7), 0.001391),
    ((4, 7, 9), 0.001374), ((0, 2, 7), 0.001346), ((4, 5, 9), 0.001341), ((0, 1, 2), 0.001338),
    ((2, 3, 5), 0.001320), ((1, 7, 9), 0.001315), ((1, 5, 7), 0.001315), ((2, 6, 8), 0.001313),
    ((1, 7, 8), 0.001309), ((0, 4, 8), 0.001266), ((0, 3, 5), 0.001260), ((1, 4, 9), 0.001258),
    ((3, 7, 9), 0.001254), ((0, 1, 7), 0.001249), ((0, 3, 7), 0.001245), ((2, 6, 7), 0.001240),
    ((0, 2, 4), 0.001238), ((4, 5, 7), 0.001235), ((5, 7, 8), 0.001220), ((0, 1, 5), 0.001203),
    ((4, 5, 8), 0.001201), ((0, 1, 3), 0.001183), ((2, 7, 8), 0.001179), ((2, 4, 6), 0.001165),
    ((2, 5, 7), 0.001165), ((0, 1, 4), 0.001163), ((3, 6, 7), 0.001153), ((0, 2, 3), 0.001152),
    ((5, 6, 8), 0.001148), ((4, 5, 6), 0.001136), ((3, 7, 8), 0.001119), ((0, 4, 7), 0.001115),
    ((6, 7, 9), 0.001113), ((3, 4, 5), 0.001111), ((2, 3, 6), 0.001110), ((1, 5, 8), 0.001107),
    ((2, 3, 7), 0.001107), ((0, 6, 7), 0.001070), ((0, 5, 8), 0.001070), ((0, 1, 6), 0.001068),
    ((0, 6, 8), 0.001067), ((0, 2, 5), 0.001058), ((2, 3, 8), 0.001057), ((5, 7, 9), 0.001055),
    ((2, 4, 7), 0.001054), ((0, 2, 8), 0.001054), ((1, 3, 7), 0.001050), ((2, 5, 8), 0.001047),
    ((3, 4, 6), 0.001047), ((0, 3, 8), 0.001038), ((0, 3, 6), 0.001034), ((2, 4, 9), 0.001031),
    ((1, 3, 9), 0.001030), ((3, 4, 9), 0.001029), ((0, 3, 9), 0.001028), ((1, 2, 9), 0.001026),
    ((2, 7, 9), 0.001016), ((1, 2, 8), 0.001008), ((1, 4, 7), 0.001007), ((1, 3, 8), 0.000986),
    ((4, 6, 9), 0.000982), ((1, 5, 9), 0.000981), ((1, 3, 5), 0.000981), ((0, 3, 4), 0.000963),
    ((0, 4, 6), 0.000960), ((3, 6, 8), 0.000932), ((2, 5, 6), 0.000931), ((0, 2, 6), 0.000926),
    ((0, 7, 9), 0.000922), ((0, 1, 8), 0.000917), ((0, 5, 6), 0.000917), ((2, 4, 5), 0.000910),
    ((1, 2, 5), 0.000899), ((3, 5, 6), 0.000896), ((1, 5, 6), 0.000883), ((0, 7, 8), 0.000881),
    ((4, 6, 8), 0.000872), ((2, 6, 9), 0.000854), ((1, 6, 8), 0.000806), ((0, 2, 9), 0.000801),
]


def matches_mpi3snp(combos, mi):
    """tests/utils.h:86-103: normalised MI within 5e-4 of MPI3SNP's."""
    norm = {c: v / MPI3SNP[0][1] for c, v in MPI3SNP}
    top = float(mi[0])
    return all(abs(float(v) / top - norm.get(tuple(int(x) for x in c), 0.0)) <= 5e-4
               for c, v in zip(combos, mi))


@pytest.fixture(scope="module")
def tped():
    return O.read_tped(TPED, TFAM)


@pytest.fixture(scope="module")
def golden_tables():
    return np.load(os.path.join(GOLDEN, "testdata_tables.npz"))


@pytest.fixture(scope="module")
def golden_search():
    with open(os.path.join(GOLDEN, "testdata_search.json")) as f:
        return json.load(f)


# --- Dataset (tests/dataset.cpp) -------------------------------------------
def test_dataset_shape(tped):
    # tests/dataset.cpp:10-23 (ALIGN == 64 branch)
    assert (tped.snps, tped.cases, tped.ctrls) == (10, 9000, 9800)
    assert (tped.cases_words, tped.ctrls_words) == (144, 160)
    # ALIGN == 32 and no-ALIGN word counts of the same test
    assert (O.lib().orc_words_for(9000, 32), O.lib().orc_words_for(9800, 32)) == (144, 156)
    assert (O.lib().orc_words_for(9000, 0), O.lib().orc_words_for(9800, 0)) == (141, 154)


def test_dataset_rows_complete(tped):
    # tests/dataset.cpp:25-44: the 3 rows of each SNP popcount to the class size
    for s in range(tped.snps):
        c, t = tped.snp(s)
        assert sum(bin(int(w)).count("1") for w in c) == 9000
        assert sum(bin(int(w)).count("1") for w in t) == 9800


def test_dataset_formats_equal(tped, golden_search):
    # tests/dataset.cpp:79-98: TPED == RAW == GAMETES RAW word for word, and
    # equal to the reference loader's tables (SHA-256 written by the reference)
    import hashlib
    raw, gam = O.read_raw(RAW), O.read_raw(GAMETES)
    assert np.array_equal(tped.words, raw.words)
    assert np.array_equal(tped.words, gam.words)
    h = hashlib.sha256(tped.words.tobytes()).hexdigest()
    assert h == golden_search["sha256"]["tped"] == golden_search["sha256"]["raw"] == golden_search["sha256"]["gametes"]


def test_pack_bit_order():
    # RAWFile.hpp:233-283: MSB-first full words, right-aligned partial word
    geno = np.zeros((1, 70), np.int8)
    geno[0, 0] = 1     # first sample: MSB of word 0 in row 1
    geno[0, 63] = 2    # 64th sample: LSB of word 0 in row 2
    geno[0, 69] = 1    # last of 6 samples in word 1: LSB of word 1 in row 1
    d = O.pack(geno, np.zeros(70, np.uint8), align=0)
    c, _ = d.snp(0)
    W = d.cases_words
    assert W == 2
    row0, row1, row2 = c[:W], c[W:2 * W], c[2 * W:]
    assert int(row1[0]) == 1 << 63 and int(row2[0]) == 1
    assert int(row1[1]) == 1 and int(row0[1]) == 0b111110
    assert int(row0[0]) == ((1 << 64) - 1) ^ (1 << 63) ^ 1


# --- GenotypeTable (tests/genotypetable.cpp) --------------------------------
def test_combine_rows(tped, golden_tables):
    # tests/genotypetable.cpp:23-36: row i = t1[i/3] & t2[i%3]
    cw, tw = tped.cases_words, tped.ctrls_words
    c0, t0 = tped.snp(0)
    c1, t1 = tped.snp(1)
    out = np.zeros(9 * cw, np.uint64)
    O.lib().orc_combine(np.ascontiguousarray(c0), 3, np.ascontiguousarray(c1), cw, out)
    for i in range(9):
        assert np.array_equal(out[i * cw:(i + 1) * cw], c0[(i // 3) * cw:(i // 3 + 1) * cw] & c1[(i % 3) * cw:(i % 3 + 1) * cw])
    ref = golden_tables["combine_0_1"]
    assert np.array_equal(out, ref[:9 * cw])


def test_popcnt_matches_bitset(tped):
    # tests/genotypetable.cpp:39-66: popcount of the combined rows, zero padding
    cw = tped.cases_words
    c0, _ = tped.snp(0)
    c1, _ = tped.snp(1)
    comb = np.zeros(9 * cw, np.uint64)
    O.lib().orc_combine(np.ascontiguousarray(c0), 3, np.ascontiguousarray(c1), cw, comb)
    ct = np.full(16, 7, np.uint32)
    O.lib().orc_popcnt(np.ascontiguousarray(c0), 3, np.ascontiguousarray(c1), cw, ct, 16)
    for i in range(9):
        assert ct[i] == sum(bin(int(w)).count("1") for w in comb[i * cw:(i + 1) * cw])
    assert (ct[9:] == 0).all()


def test_counts_match_reference_all_combinations(tped, golden_tables):
    # every combination of tests/data at orders 2-8 against the reference build
    for k in (2, 3, 4, 5, 6, 7, 8):
        combos = golden_tables[f"k{k}_combos"]
        for i, c in enumerate(combos):
            cc, ct = O.counts(tped, c)
            assert np.array_equal(cc, golden_tables[f"k{k}_cases"][i])
            assert np.array_equal(ct, golden_tables[f"k{k}_ctrls"][i])


def test_counts_bruteforce_synthetic():
    # closed form: count of samples with the cell's genotypes, incl. values
    # outside {0,1,2} (no cell) and ragged class sizes
    rng = np.random.default_rng(5)
    geno = rng.integers(0, 3, size=(7, 157)).astype(np.int8)
    geno[2, 5] = 3
    geno[4, 100] = -1
    is_ctrl = (rng.random(157) < 0.45).astype(np.uint8)
    d = O.pack(geno, is_ctrl)
    for k in (2, 3, 4, 5, 6, 7):
        for c in itertools.combinations(range(7), k):
            assert all(np.array_equal(a, b) for a, b in zip(O.counts(d, c), O.counts_bruteforce(geno, is_ctrl, c)))


# --- MutualInformation (tests/mi.cpp) ---------------------------------------
def test_mi_known_answers():
    # tests/mi.cpp:14-30 (tolerance 1e-5 absolute)
    uni = (np.ones(9, np.uint32), np.full(9, 2, np.uint32))
    one = np.zeros(9, np.uint32); one[0] = 9
    two = np.zeros(9, np.uint32); two[1] = 18
    for f in (O.mi_f32_ref512, O.mi_f32_base, O.mi_f64):
        assert abs(f(*uni, 2, 9, 18) - 0.0) < 1e-5
        assert abs(f(one, two, 2, 9, 18) - 0.6365141682948129) < 1e-5
    # the closed form of the second table: H(Y) = ln 3 - (2/3) ln 2
    assert O.mi_f64(one, two, 2, 9, 18) == pytest.approx(math.log(3) - 2 / 3 * math.log(2), abs=1e-15)


def test_mi_f32_bit_exact_vs_reference(golden_tables):
    # the reference's float MI bits, both builds, all 1,002 tables of tests/data
    for k in (2, 3, 4, 5, 6, 7, 8):
        cc, ct = golden_tables[f"k{k}_cases"], golden_tables[f"k{k}_ctrls"]
        for i in range(len(cc)):
            if O.has_avx512():
                assert np.float32(O.mi_f32_ref512(cc[i], ct[i], k, 9000, 9800)) == golden_tables[f"k{k}_mi_avx512"][i]
            assert np.float32(O.mi_f32_base(cc[i], ct[i], k, 9000, 9800)) == golden_tables[f"k{k}_mi_base"][i]


def test_mi_f64_invariants():
    rng = np.random.default_rng(9)
    for _ in range(50):
        cc = rng.integers(0, 50, 9).astype(np.uint32)
        ct = rng.integers(0, 50, 9).astype(np.uint32)
        a, b = int(cc.sum()), int(ct.sum())
        if a == 0 or b == 0:
            continue
        m = O.mi_f64(cc, ct, 2, a, b)
        assert m >= -1e-12
        # permuting the order of the two SNPs permutes the 3x3 cells
        perm = np.arange(9).reshape(3, 3).T.ravel()
        assert O.mi_f64(cc[perm], ct[perm], 2, a, b) == pytest.approx(m, abs=1e-14)
        # chi-square link: 2 N MI ~ G statistic, computed independently
        n = a + b
        g = 0.0
        for x in range(9):
            nx = int(cc[x]) + int(ct[x])
            for v, tot in ((int(cc[x]), a), (int(ct[x]), b)):
                if v:
                    g += 2 * v * math.log(v * n / (nx * tot))
        assert 2 * n * m == pytest.approx(g, rel=1e-9, abs=1e-9)
    # independence: a product table has MI 0
    cc = np.array([4, 8, 12, 2, 4, 6, 2, 4, 6], np.uint32)
    assert O.mi_f64(cc, 2 * cc, 2, int(cc.sum()), int(2 * cc.sum())) == pytest.approx(0.0, abs=1e-15)


# --- Distribution (Distribution.h) ------------------------------------------
@pytest.mark.parametrize("n,k1", [(7, 1), (7, 2), (9, 3), (6, 5)])
def test_distribution_partition(n, k1):
    full = [tuple(r) for r in O.dist_prefixes(n, k1, 1, 0)]
    assert full == list(itertools.combinations(range(n), k1))
    for step in (1, 2, 3, 5):
        got = []
        for off in range(step):
            got += [tuple(r) for r in O.dist_prefixes(n, k1, step, off)]
            # offset + j*step in lexicographic rank
            assert [tuple(r) for r in O.dist_prefixes(n, k1, step, off)] == full[off::step]
        assert sorted(got) == full


# --- Search (tests/threadedsearch.cpp) --------------------------------------
@pytest.mark.parametrize("threads", [1, 32])
def test_search_reference_test(tped, threads):
    # tests/threadedsearch.cpp:17-31
    for o in (2, 3, 4, 5, 6, 7, 8):
        combos, mi, _ = O.search(tped, o, 90, threads=threads, mi_mode=0)
        assert len(combos) > 0
        if o > 2:
            assert len(combos) == min(90, math.comb(10, o))
        assert combos.shape[1] == o
        assert len({tuple(c) for c in combos}) == len(combos)
        assert (np.diff(combos, axis=1) > 0).all()
        if o == 3:
            assert matches_mpi3snp(combos, mi)


def test_search_matches_reference_top(tped, golden_search):
    # same top-90 sets and values as ThreadedSearch(1) of the reference build
    for o in (2, 3, 4, 5, 6, 7, 8, 9):
        combos, mi, _ = O.search(tped, o, 90, threads=4, mi_mode=0)
        g = golden_search["top"][str(o)]
        assert {tuple(c) for c in combos} == {tuple(c) for c in g["combos"]}
        assert sorted(mi.tolist()) == sorted(np.float32(g["mi"]).tolist())


def test_search_synthetic_matches_reference(golden_dir):
    import fiuncho_amd as fa
    with open(os.path.join(golden_dir, "synth_search.json")) as f:
        synth = json.load(f)
    for name, g in synth.items():
        spec = fa.synth_spec(g["snps"], g["cases"], g["ctrls"], seed=g["seed"], planted=g["planted"])
        words, geno, cw, tw = fa.synth_generate(spec, words=True, geno=True)
        is_ctrl = np.r_[np.zeros(g["cases"], np.uint8), np.ones(g["ctrls"], np.uint8)]
        d = O.pack(geno, is_ctrl)
        import hashlib
        assert hashlib.sha256(d.words.tobytes()).hexdigest() == g["sha256"], name
        assert np.array_equal(d.words, words), name
        combos, mi, _ = O.search(d, g["order"], g["outputs"], threads=4, mi_mode=0)
        assert {tuple(c) for c in combos} == {tuple(c) for c in g["combos"]}, name
        assert sorted(mi.tolist()) == sorted(np.float32(g["mi"]).tolist()), name


def _write_quirk_files(tmp):
    q = np.load(os.path.join(GOLDEN, "loader_quirk.npz"))
    paths = {}
    for key, name in (("raw_text", "q.raw"), ("tped_text", "q.tped"), ("tfam_text", "q.tfam")):
        paths[name] = os.path.join(tmp, name)
        with open(paths[name], "wb") as f:
            f.write(q[key].tobytes())
    return q, paths


def test_loader_flush_quirk_matches_reference(tmp_path):
    # RAWFile.hpp:245-258 / TPEDFile.hpp:229-244: a full word flushed at the
    # 64k-th sample of a class is re-flushed as zero by the next sample of the
    # other class; the reference build's tables for such a file are stored
    q, p = _write_quirk_files(str(tmp_path))
    r = O.read_raw(p["q.raw"])
    t = O.read_tped(p["q.tped"], p["q.tfam"])
    assert [r.snps, r.cases, r.ctrls, r.cases_words, r.ctrls_words] == q["raw_shape"].tolist()
    assert np.array_equal(r.words, q["raw_words"])
    assert np.array_equal(t.words, q["tped_words"])
    # and the effect is real: the re-flushed words are zero in every row
    cw = r.cases_words
    per = 3 * (cw + r.ctrls_words)
    for s in range(r.snps):
        assert all(int(r.words[s * per + k * cw + 0]) == 0 for k in range(3))

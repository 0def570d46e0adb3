"""The weighted (ZDP) paths on the CPU: the oracle's weighted SemiGlobal and
_wighted / _wighted_1gap overloads against the golden vectors of the
compiled reference, the live reference and two closed relations, and the
library's host readers (score folder, GFF3 categories) against the
reference's Score and readGffFileWithEveryThing.

Relations used as pins (independent of the reference's code):
  * with one profile for every base and a band covering every column, the
    weighted two-piece DP scores exactly as SemiGlobal_no_avx with that
    profile's scores (they differ only in the band's last column, where no
    vertical state can enter in either);
  * the single-gap DP equals the two-piece DP whose second piece can never
    win (open2 = -1e6, ext2 = 0), cigar included.
"""
import json
import os

import numpy as np
import pytest

import dcns_amd
from dcns_amd import synth
from oracle import bind

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def fixtures():
    d = json.load(open(os.path.join(GOLD, "weighted.json")))
    d["prof"] = {k: {int(c): tuple(v) for c, v in p.items()} for k, p in d["profiles"].items()}
    return d


def test_oracle_tasks_match_golden(oracle_lib, fixtures):
    for t in fixtures["tasks"]:
        a, b = synth.encode(t["s1"]), synth.encode(t["s2"])
        prof = fixtures["prof"][t["profiles"]]
        mx, e1, e2, cig, _ = oracle_lib.semiglobal_weighted(a, b, t["wcat"], prof, w=t["w"], two_piece=t["two_piece"])
        assert [mx, e1, e2] == t["out"] and cig == t["cigar"], t
        so = oracle_lib.semiglobal_weighted(a, b, t["wcat"], prof, w=t["w"], two_piece=t["two_piece"], cigar=False)
        assert list(so[:3]) == t["out_score_only"]


def test_oracle_pairs_match_golden(oracle_lib, fixtures):
    for c in fixtures["pairs"]:
        a, b = synth.encode(c["s1"]), synth.encode(c["s2"])
        got = oracle_lib.find_similar_weighted(a, b, c["wcat"], fixtures["prof"][c["profiles"]],
                                               two_piece=c["two_piece"], **c["params"])
        assert [list(g[:5]) + [g[5], g[6], list(g[7])] for g in got] == \
            [list(w[:5]) + [w[5], w[6], list(w[7])] for w in c["fragments"]]


def test_oracle_vs_live_reference(oracle_lib, ref_lib, tmp_path, fixtures):
    prof = fixtures["prof"]["alt"]
    folder = bind.write_score_folder(str(tmp_path / "sc"), prof)
    rng = np.random.default_rng(31)
    cats = np.array(list(prof), np.int16)
    for t in range(150):
        L1, L2 = int(rng.integers(1, 120)), int(rng.integers(1, 120))
        a = rng.choice([0, 1, 2, 3, 4], L1, p=[.24, .24, .04, .24, .24]).astype(np.int8)
        b = synth.mutate(rng, a, 0.2, 0.05)[:L2] if t % 2 else rng.choice([0, 1, 3, 4], L2).astype(np.int8)
        if len(b) == 0:
            continue
        wc = rng.choice(cats, L1 + 2)
        w = int(rng.choice([2, 4, 9, 30, 20000]))
        for tp in (True, False):
            assert oracle_lib.semiglobal_weighted(a, b, wc, prof, w=w, two_piece=tp)[:4] == \
                ref_lib.semiglobal_weighted(a, b, wc, prof, w=w, two_piece=tp, folder=folder)[:4]


def test_uniform_profile_equals_semiglobal(oracle_lib):
    rng = np.random.default_rng(7)
    for t in range(80):
        sc = [(2, -3, -4, -2, -45, 0), (1, -2, -3, -1, -15, -1), (3, -4, -6, -2, -30, -1)][t % 3]
        z = int(rng.integers(5, 80))
        prof = {5: sc + (z,)}
        L1, L2 = int(rng.integers(1, 200)), int(rng.integers(1, 200))
        a = rng.choice([0, 1, 3, 4], L1).astype(np.int8)
        b = synth.mutate(rng, a, 0.15, 0.04)[:L2] if t % 3 else rng.choice([0, 1, 3, 4], L2).astype(np.int8)
        if len(b) == 0:
            continue
        w = len(a) + len(b)  # every row's band reaches column L2
        got = oracle_lib.semiglobal_weighted(a, b, np.full(L1 + 2, 5), prof, w=w, two_piece=True, cigar=False)
        want, _ = oracle_lib.semiglobal(a, b, sc, z, w, cigar=False)
        assert tuple(got[:3]) == tuple(want), (t, sc, z)


def test_single_gap_equals_two_piece_without_second_piece(oracle_lib):
    rng = np.random.default_rng(8)
    prof1 = {1: (2, -3, -4, -2, 0, 0, 40), 2: (1, -1, -2, -1, 0, 0, 25)}
    prof2 = {c: v[:4] + (-1000000, 0, v[6]) for c, v in prof1.items()}
    for t in range(80):
        L1, L2 = int(rng.integers(1, 150)), int(rng.integers(1, 150))
        a = rng.choice([0, 1, 3, 4], L1).astype(np.int8)
        b = synth.mutate(rng, a, 0.15, 0.04)[:L2] if t % 2 else rng.choice([0, 1, 3, 4], L2).astype(np.int8)
        if len(b) == 0:
            continue
        wc = rng.choice([1, 2], L1 + 2)
        w = int(rng.choice([3, 12, 500]))
        one = oracle_lib.semiglobal_weighted(a, b, wc, prof1, w=w, two_piece=False)
        two = oracle_lib.semiglobal_weighted(a, b, wc, prof2, w=w, two_piece=True)
        assert one[:4] == two[:4], t


# -- host readers of the library (no device needed) ---------------------------

def test_score_folder_matches_fixture_and_reference(tmp_path, fixtures, ref_lib):
    for name, prof in fixtures["prof"].items():
        folder = bind.write_score_folder(str(tmp_path / name), prof)
        assert dcns_amd.read_score_folder(folder) == prof
        for c in prof:
            assert ref_lib.score_profile(folder, c) == prof[c]
    # a category listed twice, and a later file with fewer lines: the later
    # file sets match / mismatch and the lines it has; the rest stay
    d = tmp_path / "odd"
    d.mkdir()
    (d / "matrixList").write_text("7\n7b\n\n9")
    (d / "7").write_text("2\n-3\n-4\n-2\n-45\n0\n50")
    (d / "7b").write_text("5\n-6\n-8")
    (d / "9").write_text("1\n-1\n-2\n-1\n-10\n-1\n30\nextra line")
    got = dcns_amd.read_score_folder(str(d))
    assert list(got) == [7, 9]
    for c in (7, 9):
        assert got[c] == ref_lib.score_profile(str(d), c), c


def test_score_folder_errors(tmp_path):
    with pytest.raises(dcns_amd.DcnsError):
        dcns_amd.read_score_folder(str(tmp_path / "missing"))
    (tmp_path / "matrixList").write_text("3\n")
    with pytest.raises(dcns_amd.DcnsError):
        dcns_amd.read_score_folder(str(tmp_path))


def test_gff_categories_match_fixture(fixtures):
    g = fixtures["gff"]
    path = os.path.join(GOLD, g["file"])
    for name, size in g["chroms"]:
        got = dcns_amd.gff_categories(path, name, size, 0, size)
        assert got.tolist() == g["categories"][name], name
        # windows past both ends read 10, as the library defines them
        w = dcns_amd.gff_categories(path, name, size, -2, 6)
        assert w[:2].tolist() == [10, 10] and w[2:].tolist() == g["categories"][name][:4]


def _random_gff(rng, chroms, path):
    kinds = ["gene", "mRNA", "exon", "five_prime_UTR", "CDS", "stop_codon", "transposable_element", "helitron",
             "repeat_region", "contig", "odd_type", "lnc_RNA"]
    lines = ["##gff-version 3"]
    for name, size in chroms:
        for _ in range(int(rng.integers(0, 40))):
            a = int(rng.integers(1, size + 1))
            b = min(size, a + int(rng.integers(0, 700)))
            if rng.random() < 0.2:
                a, b = b, a
            k = kinds[int(rng.integers(0, len(kinds)))]
            lines.append("\t".join([name, "s", k, str(a), str(b), ".", "+", ".", "ID=x"]))
        if rng.random() < 0.3:  # a chromosome with a single exon record
            lines.append("\t".join([name, "s", "exon", "10", "90", ".", "+", ".", "ID=e"]))
        lines.append("")  # blank line
        lines.append(f"{name}\ts\tgene\t5")  # too few fields: skipped
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def test_gff_categories_match_live_reference(ref_lib, tmp_path):
    rng = np.random.default_rng(11)
    for t in range(25):
        chroms = [(f"k{i}", int(rng.integers(200, 4000))) for i in range(int(rng.integers(1, 4)))]
        path = str(tmp_path / f"a{t}.gff3")
        _random_gff(rng, chroms, path)
        for k, (name, size) in enumerate(chroms):
            want = ref_lib.gff_categories(path, chroms, k, 0, size)
            got = dcns_amd.gff_categories(path, name, size, 0, size)
            assert got.tolist() == want.tolist(), (t, name)

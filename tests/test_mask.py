"""maskGenome on the CPU: the oracle against the compiled reference's golden
files and the live reference, a brute-force window check of the oracle's
scan, and the library's host readers (k-mer profile, CDS masks from GFF3 /
SAM) against the oracle.

Pins of the oracle (independent of its code):
  * golden files: the reference command's output file and stdout on 15
    inputs, and its failure on three (a sequence shorter than k, line width
    0, a CDS range past its sequence);
  * live: the same comparison on random inputs where oracle/_ref is built;
  * brute force: each window compared with every member of a list, and the
    marks set position by position.
"""
import glob
import json
import os
import shutil

import numpy as np
import pytest

import dcns_amd
from dcns_amd import synth
from oracle import bind

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mask")
CASES = sorted(glob.glob(os.path.join(GOLD, "case*")))


def _copy_case(src, dst):
    for f in os.listdir(src):
        if not f.startswith("expected") and f != "case.json":
            shutil.copy(os.path.join(src, f), dst)
    return json.load(open(os.path.join(src, "case.json")))


@pytest.mark.parametrize("case", CASES, ids=os.path.basename)
def test_oracle_matches_golden(case, tmp_path):
    meta = _copy_case(case, tmp_path)
    rc, log = bind.mask_genome(str(tmp_path), meta["opts"], "out.fa")
    if meta["status"] != 0:
        assert rc != 0
        return
    assert rc == 0, log
    assert log == meta["stdout"]
    assert open(tmp_path / "out.fa", "rb").read() == open(os.path.join(case, "expected.fa"), "rb").read()


def test_oracle_vs_live_reference(tmp_path):
    if bind.reference_mask_genome(str(tmp_path), ["-h"]) is None:
        pytest.skip("oracle/_ref (compiled reference) not built here")
    rng = np.random.default_rng(77)
    done = 0
    for t in range(40):
        case = synth.mask_case(rng)
        d = tmp_path / f"c{t}"
        d.mkdir()
        args = synth.mask_args(case, str(d), "ref.fa")
        r = bind.reference_mask_genome(str(d), args)
        rc, log = bind.mask_genome(str(d), case["opts"], "orc.fa")
        assert (r.returncode == 0) == (rc == 0), (t, r.returncode, rc, log)
        if rc == 0:
            assert r.stdout == log, t
            assert open(d / "ref.fa", "rb").read() == open(d / "orc.fa", "rb").read(), t
            done += 1
    assert done >= 30


def test_oracle_scan_brute_force(tmp_path):
    rng = np.random.default_rng(5)
    for t in range(60):
        k = int(rng.integers(1, 7))
        alphabet = list("ACGTNacg") if t % 2 else list("ACGT")
        members = ["".join(rng.choice(list("ACGTN"), k)) for _ in range(int(rng.integers(0, 12)))]
        seqs = ["".join(rng.choice(alphabet, int(rng.integers(k, 60)))) for _ in range(int(rng.integers(1, 4)))]
        for s in seqs[:2]:  # members taken from the sequences
            a = int(rng.integers(0, len(s) - k + 1))
            members.append(s[a:a + k].upper())
        dump = tmp_path / f"d{t}.fa"
        recs = [(m, 10) for m in members] + [("ACGTACGTAC"[:k], 1)]
        dump.write_text("".join(f">{c}\n{m}\n" for m, c in recs))
        k_o, masks = bind.kmer_mask(str(dump), 5, [s.encode() for s in seqs])
        assert k_o == k
        for s, m in zip(seqs, masks):
            want = np.zeros(len(s), np.uint8)
            for i in range(len(s) - k + 1):
                if any(s[i:i + k] == x for x in members):  # case-sensitive, as read
                    want[i:i + k] = 1
            assert m.tolist() == want.tolist(), (t, s, members)


# -- host readers of the library ----------------------------------------------

def _members_product(path, freq):
    ks = dcns_amd.KmerSet(path, freq)
    try:
        return ks.k, sorted(set(ks.kmers())), ks.members
    finally:
        ks.close()


def test_kmer_profile_matches_oracle(tmp_path):
    paths = [os.path.join(c, "kmers.fa") for c in CASES if os.path.exists(os.path.join(c, "kmers.fa"))]
    rng = np.random.default_rng(9)
    for t in range(30):
        case = synth.mask_case(rng)
        p = tmp_path / f"k{t}.fa"
        p.write_text(case["dump"])
        paths.append(str(p))
    # odd dumps: no header, CRLF lines, trailing text without newline, counts with sign / text after them
    odd = {"noheader.fa": "ACGT\nACGA\n", "crlf.fa": ">9\r\nACGT\r\n>9\r\nACGA\r\n",
           "nonl.fa": ">9\nACGT\n>9\nACG", "sign.fa": ">+9 x\nacgt\n>-3\nACGA\n>\t12\nAAAA\n>7abc\nTTTT\n",
           "empty.fa": "", "mixed.fa": ">9\nACGTA\n>9\nAC\n>9\nNNNNN\n>9\nAcGtT\n"}
    for f, text in odd.items():
        (tmp_path / f).write_text(text)
        paths.append(str(tmp_path / f))
    for p in paths:
        for freq in (1, 3, 6):
            k_o, want = bind.kmer_members(p, freq)
            k_p, got, n = _members_product(p, freq)
            assert (k_p, got) == (k_o, want), (p, freq)
            assert n >= len(want)
    assert _members_product(paths[0], 0)[:2] == (10, [])  # no positive threshold: nothing read


def test_kmer_profile_errors(tmp_path):
    with pytest.raises(dcns_amd.DcnsError):
        dcns_amd.KmerSet(str(tmp_path / "missing.fa"), 5)
    (tmp_path / "bad.fa").write_text(">x9\nACGT\n")
    with pytest.raises(dcns_amd.DcnsError):
        dcns_amd.KmerSet(str(tmp_path / "bad.fa"), 5)


def _oracle_cds(d, case):
    """The oracle's CDS masks of a case: its command run with no k-mers,
    soft masking, no line breaks, over a genome of the same names and sizes
    made of 'A' only (a masked position is then 'a')."""
    with open(os.path.join(d, "genome.fa"), "w") as f:
        for n, L in case["sizes"]:
            f.write(f">{n}\n{'A' * L}\n")
    o = dict(case["opts"], use_dump=False, soft=True, width=-1)
    rc, log = bind.mask_genome(d, o, "cds_out.fa")
    if rc != 0:
        return None, log
    lines = open(os.path.join(d, "cds_out.fa")).read().split("\n")
    masks = {}
    for i in range(0, len(lines) - 1, 2):
        masks[lines[i][1:]] = np.frombuffer(lines[i + 1].encode(), np.uint8) == ord("a")
    return masks, log


def test_cds_masks_match_oracle(tmp_path):
    rng = np.random.default_rng(13)
    checked = 0
    for t in range(60):
        case = synth.mask_case(rng)
        case["opts"]["use_gff"] = t % 3 != 2
        case["opts"]["use_sam"] = t % 3 != 0
        d = tmp_path / f"c{t}"
        d.mkdir()
        synth.mask_args(case, str(d))
        sizes = [(n, int(rng.integers(1, 3000))) for n in dict.fromkeys(case["names"])]
        case["sizes"] = sizes
        want, log = _oracle_cds(str(d), case)
        names, lens = [n for n, _ in sizes], [L for _, L in sizes]
        m = dcns_amd.CdsMask(names, lens)
        try:
            if case["opts"]["use_gff"]:
                m.gff(str(d / "ann.gff3"), case["opts"]["gene"])
            rebuilt = m.sam(str(d / "cds.sam"), str(d / "cds.fa"), case["opts"]["similarity"]) \
                if case["opts"]["use_sam"] else False
        except dcns_amd.DcnsError:
            assert want is None, (t, log)
            continue
        assert want is not None, (t, log)
        assert rebuilt == ("GFF was not used for genome masking" in log), t
        for i, (n, L) in enumerate(sizes):
            assert m.mask(i).astype(bool).tolist() == want[n].tolist(), (t, n)
        m.close()
        checked += 1
    assert checked >= 40


def test_cds_mask_errors(tmp_path):
    m = dcns_amd.CdsMask(["a"], [20])
    with pytest.raises(dcns_amd.DcnsError):
        m.gff(str(tmp_path / "missing.gff3"))
    (tmp_path / "past.gff3").write_text("a\ts\tCDS\t30\t40\t.\t+\t.\tID=x\n")
    with pytest.raises(dcns_amd.DcnsError):
        m.gff(str(tmp_path / "past.gff3"))
    (tmp_path / "grow.gff3").write_text("a\ts\tCDS\t15\t40\t.\t+\t.\tID=x\na\ts\tCDS\t35\t41\t.\t+\t.\tID=y\n")
    m.gff(str(tmp_path / "grow.gff3"))  # the first range grows the mask string to 40: the second starts inside
    assert m.mask(0).tolist() == [0] * 14 + [1] * 6
    with pytest.raises(dcns_amd.DcnsError):
        dcns_amd.CdsMask(["a", "a"], [3, 4])
    assert dcns_amd.mask_body_size(120, 60) == 123 and dcns_amd.mask_body_size(7, -1) == 8
    assert dcns_amd.mask_body_size(0, 60) == 1

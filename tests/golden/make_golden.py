"""Writes the golden fixtures in this folder from the compiled reference.

Run in a container where the reference sources are present:

    make -C oracle ref && python tests/golden/make_golden.py

Every fixture holds inputs (sequences as A/C/N/G/T text, parameters) and the
outputs of the reference build (oracle/_ref) on them: K1 SmithWaterman
(striped and textbook paths), SemiGlobal_xextend with and without traceback,
the two-piece SemiGlobal, the seed set before/after link_seeds, the fragments
of findSimilarFragmentsForPairedSequence (overloads 1 and 4), the weighted
paths (score profiles, weighted SemiGlobal tasks, the _wighted / _wighted_1gap
overloads, GFF categories), overload 3 (seeded), and the text outputs of the
cut1Gap, cut2Gaps, pairCnsXExtend, ranSco, weighted1Gap, weighted2Gaps and
cut2Gaps2 commands, and the maskGenome command's files.  Only data is
written.  ``--weighted`` / ``--seeded`` / ``--mask`` write those fixtures
only; ``--batch`` the cut1Gap command-file fixture.
"""
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_ROOT = os.environ.get("DCNS_REFERENCE", "/root/reference")
sys.path.insert(0, ROOT)

from dcns_amd import synth  # noqa: E402
from oracle import bind  # noqa: E402


def txt(a):
    return synth.to_text(a)


def main():
    ref = bind.reference()
    if ref is None:
        sys.exit("oracle/_ref is not built")
    rng = np.random.default_rng(20261203)
    scorings = [(2, -3, -4, -2), (1, -1, -2, -1), (3, -4, -6, -3), (1, -4, -2, -3), (2, -1, -2, -1)]

    # K1
    k1 = []
    for t in range(160):
        sc = scorings[t % len(scorings)]
        L1, L2 = int(rng.integers(0, 90)), int(rng.integers(0, 50))
        a = rng.choice([0, 1, 2, 3, 4], L1, p=[.24, .24, .04, .24, .24]).astype(np.int8)
        b = rng.choice([0, 1, 2, 3, 4], L2, p=[.24, .24, .04, .24, .24]).astype(np.int8)
        if L1 > 20 and L2 > 10 and t % 2:
            k = int(rng.integers(0, L1 - 10))
            m = synth.mutate(rng, a[k:k + L2], 0.1, 0.05)[:L2]
            b[:len(m)] = m
        out = list(ref.sw(a, b, sc))
        if out[0] == 0:
            out[1] = out[2] = -1  # positions are not written by the reference when nothing scores
        k1.append({"ref": txt(a), "read": txt(b), "sc": sc, "textbook": False, "out": out})
        k1.append({"ref": txt(a), "read": txt(b), "sc": sc, "textbook": True, "out": list(ref.sw(a, b, sc, True))})
    # long pairs past the int16 bound (textbook path inside SmithWaterman)
    a = synth.random_seq(rng, 11500)
    b = synth.mutate(rng, a, 0.05, 0.01)[:11200]
    k1.append({"ref": txt(a), "read": txt(b), "sc": (1, -2, -5, -3), "textbook": False,
               "out": list(ref.sw(a, b, (1, -2, -5, -3)))})
    json.dump(k1, open(os.path.join(HERE, "k1_smithwaterman.json"), "w"))

    # K2 / K3 / K4
    ext = []
    for t in range(120):
        sc = scorings[t % len(scorings)]
        L1, L2 = int(rng.integers(1, 260)), int(rng.integers(1, 260))
        a = rng.choice([0, 1, 3, 4], L1).astype(np.int8)
        b = (synth.mutate(rng, a, 0.15, 0.04)[:L2] if t % 4 else rng.choice([0, 1, 3, 4], L2).astype(np.int8))
        if len(b) == 0:
            continue
        w, xd = int(rng.integers(1, 15)), int(rng.integers(5, 40))
        sc2 = sc + (-int(rng.integers(10, 50)), -int(rng.integers(0, 2)))
        z, w2 = int(rng.integers(10, 80)), int(rng.integers(3, 400))
        tb, cig = ref.xextend_tb(a, b, sc, xd, w)
        sg, sgc = ref.semiglobal(a, b, sc2, z, w2)
        ext.append({"s1": txt(a), "s2": txt(b), "sc": sc, "w": w, "xdrop": xd,
                    "xextend": list(ref.xextend(a, b, sc, xd, w)), "xextend_tb": list(tb), "cigar": cig,
                    "sc2": sc2, "zdrop": z, "bandwidth": w2, "semiglobal": list(sg), "semiglobal_cigar": sgc})
    json.dump(ext, open(os.path.join(HERE, "k234_extension.json"), "w"))

    # seeds and fragments
    pipe = []
    for t in range(10):
        L1 = int(rng.integers(300, 3000))
        L2 = int(rng.integers(100, 3000))
        a, b = synth.one_pair(seed=7000 + t, L1=L1, L2=L2, nfrag=int(rng.integers(0, 12)))
        pipe.append({"s1": txt(a), "s2": txt(b), "seeds": ref.seed_set(a, b), "linked": ref.seed_set(a, b, linked=True),
                     "fragments": ref.find_similar(a, b)})
    a, b = synth.one_pair()
    pipe.append({"s1": txt(a), "s2": txt(b), "seeds": ref.seed_set(a, b), "linked": ref.seed_set(a, b, linked=True),
                 "fragments": ref.find_similar(a, b), "note": "10 kb x 10 kb, 20 planted fragments"})
    json.dump(pipe, open(os.path.join(HERE, "pipeline_fragments.json"), "w"))

    # cut1Gap command: inputs and the reference's text outputs
    d = os.path.join(HERE, "cut1gap")
    shutil.rmtree(d, ignore_errors=True)
    os.makedirs(d)
    write_cut1gap_inputs(d, rng)
    for f in os.listdir(d):
        if f.startswith("out"):
            os.remove(os.path.join(d, f))
    for inp, out in (("in.fa", "out"), ("in_refminus.fa", "outm")):
        r = subprocess.run([bind.REF_CLI, "cut1Gap", "-i", inp, "-r", "ref", "-ra", "refgenome.fa", "-qa",
                            "querygenome.fa", "-o", out], cwd=d, capture_output=True, text=True)
        open(os.path.join(d, f"expected_stdout_{out}.txt"), "w").write(r.stdout)
    for f in sorted(os.listdir(d)):
        if f.startswith("out"):
            os.replace(os.path.join(d, f), os.path.join(d, "expected_" + f))

    # ranSco command
    d = os.path.join(HERE, "ransco")
    shutil.rmtree(d, ignore_errors=True)
    os.makedirs(d)
    for name, n in (("refgenome.fa", 3), ("q1.fa", 2), ("q2.fa", 2)):
        with open(os.path.join(d, name), "w") as fh:
            for k in range(n):
                s = synth.to_text(synth.random_seq(rng, int(rng.integers(3000, 9000)), float(rng.uniform(.35, .6))))
                s = list(s)
                for _ in range(3):  # a few n / N runs that the sampler has to reject or strip
                    p = int(rng.integers(0, len(s) - 200))
                    ch = "n" if rng.random() < .5 else "N"
                    ln = int(rng.integers(5, 150))
                    s[p:p + ln] = [ch] * ln
                s = "".join(s)
                fh.write(f">{name[:-3]}_{k}\n")
                for i in range(0, len(s), 60):
                    fh.write(s[i:i + 60] + "\n")
    for args in (["-l", "300", "-p", "200", "-s", "5"], ["-l", "500", "-p", "100", "-s", "11", "-A", "1", "-B", "-2",
                                                       "-O", "-3", "-E", "-1"]):
        r = subprocess.run([bind.REF_CLI, "ranSco", "-r", "refgenome.fa", "-i", "q1.fa", "q2.fa"] + args, cwd=d,
                           capture_output=True, text=True)
        tag = "_".join(a.strip("-") for a in args)
        open(os.path.join(d, f"expected_{tag}.txt"), "w").write(r.stdout)
    write_two_piece(HERE, ref)
    write_weighted(HERE, ref)
    write_seeded(HERE, ref)
    print("golden fixtures written to", HERE)


TWO_PIECE_SETTINGS = [
    dict(sc=(2, -3, -4, -2, -45, 0), zdrop=50, bandwidth=20000),
    dict(sc=(2, -3, -4, -2, -45, 0), zdrop=50, bandwidth=150),
    dict(sc=(2, -3, -4, -2, -20, -1), zdrop=30, bandwidth=20000),
    dict(sc=(1, -2, -3, -1, -15, -1), zdrop=20, bandwidth=64, min_score=20),
    dict(sc=(2, -3, -4, -2, -45, 0), zdrop=10, bandwidth=3),
]


def write_two_piece(here, ref):
    """Overload 4 (the two-piece path) on planted pairs under several
    settings, and the cut2Gaps / pairCns2Gaps commands."""
    rng = np.random.default_rng(20261204)
    two = []
    for t in range(10):
        kw = TWO_PIECE_SETTINGS[t % len(TWO_PIECE_SETTINGS)]
        L1, L2 = int(rng.integers(600, 3500)), int(rng.integers(300, 3000))
        a, b = synth.one_pair(seed=9000 + t, L1=L1, L2=L2, nfrag=int(rng.integers(1, 10)))
        two.append({"s1": txt(a), "s2": txt(b), "params": kw, "fragments": ref.find_similar_2gaps(a, b, **kw)})
    json.dump(two, open(os.path.join(here, "two_piece_fragments.json"), "w"))
    d = os.path.join(here, "cut2gaps")
    shutil.rmtree(d, ignore_errors=True)
    os.makedirs(d)
    write_cut1gap_inputs(d, rng)
    write_pair_inputs(d, rng)
    # the reference's pairCns2Gaps aborts on every input with a query (its
    # query sequences are cleared before use), so it has no fixture here
    runs = [("cut2Gaps", ["-i", "in.fa", "-r", "ref", "-ra", "refgenome.fa", "-qa", "querygenome.fa", "-o", "c2_"]),
            ("cut2Gaps", ["-i", "in_refminus.fa", "-r", "ref", "-ra", "refgenome.fa", "-qa", "querygenome.fa",
                          "-o", "c2m_", "-O2", "-30", "-E2", "-1", "-u2", "500"]),
            ("pairCnsXExtend", ["-i", "pair.fa", "-r", "ref", "-o", "px_"]),
            ("pairCnsXExtend", ["-i", "pair_refminus.fa", "-r", "ref", "-o", "pxm_", "-O", "-3", "-E", "-1"])]
    for cmd, args in runs:
        r = subprocess.run([bind.REF_CLI, cmd] + args, cwd=d, capture_output=True, text=True)
        tag = args[args.index("-o") + 1]
        open(os.path.join(d, f"expected_stdout_{tag}.txt"), "w").write(r.stdout)
    for f in sorted(os.listdir(d)):
        if f.startswith(("c2_", "c2m_", "px_", "pxm_")):
            os.replace(os.path.join(d, f), os.path.join(d, "expected_" + f))


# the reference's scoreMatrix folder (scoreMatrix/0 .. 100), one profile per
# category: match, mismatch, open1, ext1, open2, ext2, zdrop
REF_SCORE_DIR = os.path.join(REF_ROOT, "scoreMatrix")
# a second set: smaller scores, a non-zero second extend (the z-drop's l term)
ALT_PROFILES = {0: (0, 0, 0, 0, 0, 0, 1073741824), 1: (1, -1, -2, -1, -15, -1, 40), 10: (2, -3, -4, -2, -30, -1, 60),
                70: (3, -4, -6, -3, -40, -1, 80), 85: (3, -4, -6, -2, -40, 0, 90), 100: (4, -5, -8, -3, -50, -1, 120)}


def category_runs(rng, n, cats):
    """Category array of n values made of runs (genes, exons, repeats)."""
    wc = np.full(n, 10, np.int16)
    for _ in range(max(1, n // 300)):
        s = int(rng.integers(0, n))
        e = min(n, s + int(rng.integers(20, 900)))
        wc[s:e] = rng.choice(cats)
    return wc


def write_weighted(here, ref):
    """The weighted (ZDP) paths: score profiles, single DP tasks, overloads
    _wighted / _wighted_1gap, GFF categories, and the weighted1Gap /
    weighted2Gaps commands."""
    rng = np.random.default_rng(20261205)
    cats = [0, 1, 10, 70, 85, 100]
    default = {c: ref.score_profile(REF_SCORE_DIR, c) for c in cats}
    alt_dir = bind.write_score_folder(tempfile.mkdtemp(), ALT_PROFILES)
    sets = {"default": (default, REF_SCORE_DIR), "alt": (ALT_PROFILES, alt_dir)}
    out = {"profiles": {k: {str(c): list(v) for c, v in p.items()} for k, (p, _) in sets.items()}, "tasks": [],
           "pairs": []}
    for t in range(96):
        name = "default" if t % 2 else "alt"
        prof, folder = sets[name]
        L1, L2 = int(rng.integers(1, 300)), int(rng.integers(1, 300))
        a = rng.choice([0, 1, 2, 3, 4], L1, p=[.24, .24, .04, .24, .24]).astype(np.int8)
        b = (synth.mutate(rng, a, 0.15, 0.04)[:L2] if t % 4 else
             rng.choice([0, 1, 3, 4], L2).astype(np.int8))
        if len(b) == 0:
            continue
        wc = category_runs(rng, L1 + 2, cats)
        w = int(rng.choice([2, 3, 7, 40, 300, 20000]))
        tp = bool(t % 3)
        mx, e1, e2, cig, _ = ref.semiglobal_weighted(a, b, wc, prof, w=w, two_piece=tp, folder=folder)
        mx2, f1, f2, _, _ = ref.semiglobal_weighted(a, b, wc, prof, w=w, two_piece=tp, cigar=False, folder=folder)
        out["tasks"].append({"s1": txt(a), "s2": txt(b), "wcat": [int(x) for x in wc], "profiles": name, "w": w,
                             "two_piece": tp, "out": [mx, e1, e2], "cigar": cig, "out_score_only": [mx2, f1, f2]})
    for t in range(8):
        name = "default" if t % 2 else "alt"
        prof, folder = sets[name]
        L1, L2 = int(rng.integers(800, 3500)), int(rng.integers(500, 3000))
        a, b = synth.one_pair(seed=9500 + t, L1=L1, L2=L2, nfrag=int(rng.integers(2, 12)))
        wc = category_runs(rng, L1 + 2, cats)
        for tp in (True, False):
            kw = dict(bandwidth=int(rng.choice([20000, 300, 40])))
            out["pairs"].append({"s1": txt(a), "s2": txt(b), "wcat": [int(x) for x in wc], "profiles": name,
                                 "two_piece": tp, "params": kw,
                                 "fragments": ref.find_similar_weighted(a, b, wc, prof, two_piece=tp, folder=folder,
                                                                        **kw)})
    # GFF categories of a synthetic annotation
    chroms = [("c1", 30000), ("c2", 12000)]
    gff = write_gff(os.path.join(here, "weighted_annotation.gff3"), rng, chroms)
    out["gff"] = {"file": "weighted_annotation.gff3", "chroms": chroms,
                  "categories": {n: [int(x) for x in ref.gff_categories(gff, chroms, k, 0, size)]
                                 for k, (n, size) in enumerate(chroms)}}
    json.dump(out, open(os.path.join(here, "weighted.json"), "w"))
    write_weighted_commands(here, rng, default)


def write_gff(path, rng, chroms):
    """A GFF3 with every feature class the category painter knows, nested and
    overlapping records, duplicates, reversed coordinates, comments, an
    ignored type and one unknown type."""
    types_gene = ["gene", "mRNA", "transcript", "ncRNA_gene", "lnc_RNA"]
    types_exon = ["exon", "five_prime_UTR", "three_prime_utr"]
    types_cds = ["CDS", "start_codon", "stop_codon"]
    types_te = ["transposable_element", "LTR_retrotransposon", "helitron", "repeat_region"]
    lines = ["##gff-version 3"]
    for name, size in chroms:
        lines.append(f"{name}\tsrc\tchromosome\t1\t{size}\t.\t.\t.\tID={name}")
        pos = int(rng.integers(100, 900))
        while pos < size - 3000:
            glen = int(rng.integers(800, 2800))
            g0, g1 = pos, pos + glen
            lines.append(f"{name}\tsrc\t{rng.choice(types_gene)}\t{g0}\t{g1}\t.\t+\t.\tID=g{pos}")
            lines.append(f"{name}\tsrc\tmRNA\t{g0}\t{g1}\t.\t+\t.\tID=t{pos}")  # duplicate extent
            e = g0
            while e < g1 - 100:
                el = int(rng.integers(60, 400))
                ee = min(g1, e + el)
                a_, b_ = (e, ee) if rng.random() < 0.9 else (ee, e)  # some reversed coordinates
                lines.append(f"{name}\tsrc\t{rng.choice(types_exon)}\t{a_}\t{b_}\t.\t+\t.\tParent=t{pos}")
                if rng.random() < 0.7:
                    c0 = e + int(rng.integers(0, 30))
                    lines.append(f"{name}\tsrc\t{rng.choice(types_cds)}\t{c0}\t{max(c0, ee - 20)}\t.\t+\t0\tParent=t{pos}")
                if rng.random() < 0.2:  # overlapping exon record
                    lines.append(f"{name}\tsrc\texon\t{e + 10}\t{ee + 50}\t.\t+\t.\tParent=t{pos}")
                e = ee + int(rng.integers(50, 300))
            if rng.random() < 0.5:
                t0 = g1 + int(rng.integers(50, 500))
                lines.append(f"{name}\tsrc\t{rng.choice(types_te)}\t{t0}\t{t0 + int(rng.integers(100, 900))}\t.\t.\t.\tID=te{pos}")
            pos = g1 + int(rng.integers(300, 2500))
        lines.append(f"# comment line on {name}")
    lines.append(f"{chroms[0][0]}\tsrc\tcontig\t1\t500\t.\t.\t.\tID=ctg")
    lines.append(f"{chroms[0][0]}\tsrc\tmystery_feature\t100\t200\t.\t.\t.\tID=x")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return path


def write_weighted_commands(here, rng, default):
    """weighted1Gap / weighted2Gaps on a reference region annotated by a GFF,
    with the reference's own scoreMatrix folder and the alternative one."""
    d = os.path.join(here, "weighted_cmd")
    shutil.rmtree(d, ignore_errors=True)
    os.makedirs(d)
    chrom = synth.random_seq(rng, 16000)
    chroms = [("cr", len(chrom)), ("cz", 5000)]
    fasta_with_fai(os.path.join(d, "refgenome.fa"), [("cr", synth.to_text(chrom)),
                                                     ("cz", synth.to_text(synth.random_seq(rng, 5000)))])
    write_gff(os.path.join(d, "ann.gff3"), rng, chroms)
    bind.write_score_folder(os.path.join(d, "alt_scores"), ALT_PROFILES)
    r0, r1 = 3001, 7400  # the reference region (1-based, inclusive)
    region = chrom[r0 - 1:r1]
    qs = []
    for k in range(3):
        q = synth.random_seq(rng, int(rng.integers(1500, 3200)))
        for _ in range(5):
            ln = int(rng.integers(40, 220))
            p1 = int(rng.integers(0, len(region) - ln))
            p2 = int(rng.integers(0, len(q) - ln))
            m = synth.mutate(rng, region[p1:p1 + ln], 0.12, 0.02)[:ln]
            q[p2:p2 + len(m)] = m
        qs.append(q)
    with open(os.path.join(d, "in.fa"), "w") as fh:
        fh.write(f">ref\tspecies:spA\tchr:cr\tstrand:1\tstart:{r0}\tend:{r1}\n{synth.to_text(region)}\n")
        for k, (q, st) in enumerate(zip(qs, (1, 0, 1))):
            fh.write(f">q{k}\tspecies:sp{'BCD'[k]}\tchr:cq{k}\tstrand:{st}\tstart:{21 + k}\tend:{20 + k + len(q)}\n")
            fh.write(synth.to_text(q) + "\n")
    bind.write_score_folder(os.path.join(d, "ref_scores"), default)
    runs = [("weighted1Gap", ["-sp", "ref_scores", "-o", "w1_"]),
            ("weighted2Gaps", ["-sp", "ref_scores", "-o", "w2_"]),
            ("weighted1Gap", ["-sp", "alt_scores", "-o", "w1a_", "-O1", "-3", "-E1", "-1"]),
            ("weighted2Gaps", ["-sp", "alt_scores", "-o", "w2a_", "-u2", "300"])]
    for cmd, args in runs:
        full = ["-i", "in.fa", "-r", "ref", "-ra", "refgenome.fa", "-f", "ann.gff3"] + args
        r = subprocess.run([bind.REF_CLI, cmd] + full, cwd=d, capture_output=True, text=True)
        tag = args[args.index("-o") + 1]
        open(os.path.join(d, f"expected_stdout_{tag}.txt"), "w").write(r.stdout)
    for f in sorted(os.listdir(d)):
        if f.startswith(("w1_", "w2_", "w1a_", "w2a_")):
            os.replace(os.path.join(d, f), os.path.join(d, "expected_" + f))


def write_seeded(here, ref):
    """Overload 3 (cut2Gaps2) on planted pairs with seeds taken from the
    reference's own fragments (some ends moved off a match, which triggers the
    reference's end correction), and the cut2Gaps2 command on the cut1Gap
    fixture genomes with the cut1Gap SAM outputs as input."""
    rng = np.random.default_rng(20261206)
    cases = []
    for t in range(8):
        L1, L2 = int(rng.integers(900, 3500)), int(rng.integers(900, 3500))
        a, b = synth.one_pair(seed=9700 + t, L1=L1, L2=L2, nfrag=int(rng.integers(3, 10)))[:2]
        seeds = []
        for f in ref.find_similar(a, b):
            d = int(rng.integers(-2, 3))
            seeds.append([f[0] - 1, min(f[1] - 1 + d, L1 - 1), f[2] - 1, min(f[3] - 1 + d, L2 - 1)])
        for _ in range(3):
            e1, e2 = int(rng.integers(0, L1)), int(rng.integers(0, L2))
            seeds.append([max(0, e1 - 40), e1, max(0, e2 - 40), e2])
        if seeds:
            seeds.append(list(seeds[0]))
        kw = [dict(max_len=140000), dict(max_len=400), dict(max_len=140000, bandwidth=60),
              dict(max_len=140000, sc=(2, -3, -4, -2, -20, -1), zdrop=30)][t % 4]
        cases.append({"s1": txt(a), "s2": txt(b), "seeds": seeds, "params": kw,
                      "fragments": ref.find_similar_seeded(a, b, seeds, **kw)})
    json.dump(cases, open(os.path.join(here, "seeded_fragments.json"), "w"))
    d = os.path.join(here, "cut2gaps2")
    shutil.rmtree(d, ignore_errors=True)
    os.makedirs(d)
    src = os.path.join(here, "cut1gap")
    for f in ("refgenome.fa", "querygenome.fa"):
        shutil.copy(os.path.join(src, f), d)
    with open(os.path.join(d, "in.sam"), "w") as fh:  # every cut1Gap alignment, both input files
        for f in sorted(os.listdir(src)):
            if f.startswith("expected_out") and f.endswith(".sam"):
                fh.write(open(os.path.join(src, f)).read())
    runs = [["-o", "g2_", "-m", "12000"],
            ["-o", "g2s_", "-m", "3000", "-u2", "500", "-O2", "-30", "-E2", "-1"]]
    for args in runs:
        full = ["-sa", "in.sam", "-ra", "refgenome.fa", "-qa", "querygenome.fa"] + args
        r = subprocess.run([bind.REF_CLI, "cut2Gaps2"] + full, cwd=d, capture_output=True, text=True)
        tag = args[1]
        open(os.path.join(d, f"expected_stdout_{tag}.txt"), "w").write(r.stdout)
    for f in sorted(os.listdir(d)):
        if f.startswith(("g2_", "g2s_")):
            os.replace(os.path.join(d, f), os.path.join(d, "expected_" + f))


def fasta_with_fai(path, records, width=60):
    """Writes a FASTA file and its samtools-style .fai index."""
    with open(path, "w") as fh, open(path + ".fai", "w") as fai:
        off = 0
        for name, seq in records:
            head = f">{name}\n"
            fh.write(head)
            off += len(head)
            fai.write(f"{name}\t{len(seq)}\t{off}\t{width}\t{width + 1}\n")
            for i in range(0, len(seq), width):
                line = seq[i:i + width] + "\n"
                fh.write(line)
                off += len(line)


def write_pair_inputs(d, rng):
    """pairCns inputs: the reference region and three query regions written
    inline (with region metadata), queries carrying mutated copies of
    reference stretches; a second file with the reference on strand 0."""
    a = synth.random_seq(rng, 2400)
    qs = []
    for k in range(3):
        q = synth.random_seq(rng, int(rng.integers(1200, 2600)))
        for _ in range(4):
            ln = int(rng.integers(50, 200))
            p1 = int(rng.integers(0, len(a) - ln))
            p2 = int(rng.integers(0, len(q) - ln))
            m = synth.mutate(rng, a[p1:p1 + ln], 0.1, 0.02)[:ln]
            q[p2:p2 + len(m)] = m
        qs.append(q)
    for fname, rstrand in (("pair.fa", 1), ("pair_refminus.fa", 0)):
        with open(os.path.join(d, fname), "w") as fh:
            fh.write(f">ref\tspecies:spA\tchr:cr\tstrand:{rstrand}\tstart:101\tend:2500\n{synth.to_text(a)}\n")
            for k, (q, st) in enumerate(zip(qs, (1, 0, 1))):
                fh.write(f">q{k}\tspecies:sp{'BCD'[k]}\tchr:cq{k}\tstrand:{st}\tstart:{11 + k}\tend:{10 + k + len(q)}\n")
                fh.write(synth.to_text(q) + "\n")


def write_cut1gap_inputs(d, rng):
    """Two small genomes with masked (lowercase n) stretches, lowercase bases,
    and an -i file of one reference region and four query regions on both
    strands."""
    def masked(codes, nruns):
        s = list(synth.to_text(codes))
        for _ in range(nruns):
            p = int(rng.integers(0, len(s) - 100))
            ln = int(rng.integers(5, 90))
            s[p:p + ln] = "n" * len(s[p:p + ln])
        for _ in range(nruns):  # soft-masked lowercase bases
            p = int(rng.integers(0, len(s) - 50))
            s[p:p + 30] = [c.lower() if c != "n" else c for c in s[p:p + 30]]
        return "".join(s)

    r_chr = synth.random_seq(rng, 9000)
    q_chr1 = synth.random_seq(rng, 7000)
    q_chr2 = synth.random_seq(rng, 6000)
    # plant conserved fragments between the reference region and query regions
    for q, off in ((q_chr1, 500), (q_chr2, 300)):
        for _ in range(10):
            ln = int(rng.integers(40, 160))
            p1 = int(rng.integers(1000, 6000 - ln))
            p2 = int(rng.integers(off, min(len(q), off + 4000) - ln))
            m = synth.mutate(rng, r_chr[p1:p1 + ln], 0.1, 0.02)[:ln]
            q[p2:p2 + len(m)] = m
    # a reverse-strand copy for the strand-0 cases
    rc = (4 - r_chr[2000:2200][::-1]).astype(np.int8)
    q_chr2[4500:4700] = rc
    fasta_with_fai(os.path.join(d, "refgenome.fa"), [("chr1", masked(r_chr, 12))])
    fasta_with_fai(os.path.join(d, "querygenome.fa"), [("qc1", masked(q_chr1, 10)), ("qc2", masked(q_chr2, 10))])
    with open(os.path.join(d, "in.fa"), "w") as fh:
        regions = [("ref", "spA", "chr1", 1, 1001, 6000), ("q1", "spB", "qc1", 1, 401, 4600),
                   ("q2", "spC", "qc2", 1, 201, 4400), ("q3", "spD", "qc2", 0, 3001, 6000),
                   ("q4", "spE", "qc1", 0, 101, 3000)]
        for name, sp, chrom, strand, st, en in regions:
            fh.write(f">{name}\tspecies:{sp}\tchr:{chrom}\tstrand:{strand}\tstart:{st}\tend:{en}\n")
            fh.write("ACGT\n")
    with open(os.path.join(d, "in_refminus.fa"), "w") as fh:
        regions = [("ref", "spA", "chr1", 0, 1001, 6000), ("q1", "spB", "qc1", 1, 401, 4600),
                   ("q3", "spD", "qc2", 0, 3001, 6000)]
        for name, sp, chrom, strand, st, en in regions:
            fh.write(f">{name}\tspecies:{sp}\tchr:{chrom}\tstrand:{strand}\tstart:{st}\tend:{en}\n")
            fh.write("ACGT\n")


def write_mask(here):
    """maskGenome: inputs and the reference command's output file and stdout,
    one folder per case under mask/ (case.json: flags, expected status)."""
    root = os.path.join(here, "mask")
    if os.path.isdir(root):
        shutil.rmtree(root)
    os.makedirs(root)
    rng = np.random.default_rng(20261204)
    cases = []
    while len(cases) < 14:  # random inputs that the reference completes on
        c = synth.mask_case(rng)
        with tempfile.TemporaryDirectory() as d:
            args = synth.mask_args(c, d)
            r = bind.reference_mask_genome(d, args)
            if r.returncode == 0:
                cases.append(c)
    # a realistic one: a repetitive chromosome pair and its 20-mer counts
    g1 = synth.repeat_genome(rng, 150000, families=25)
    g2 = synth.repeat_genome(rng, 60000, families=10)
    km, cnt = synth.kmer_counts(np.concatenate([g1, g2]), 20, 1)
    sel = rng.random(len(km)) < 0.5  # part of the low-count tail
    cases.append({"genome": f">chr1\n{synth._wrap(g1.tobytes().decode(), 60)}\n>chr2\n"
                            f"{synth._wrap(g2.tobytes().decode(), 60)}\n",
                  "dump": synth.dump_text(km[sel], cnt[sel]), "gff": "", "sam": "", "cds": "",
                  "opts": {"width": 60, "soft": False, "alt": "n", "gene": False, "similarity": 0.8, "freq": 3,
                           "use_dump": True, "use_gff": False, "use_sam": False}, "k": 20, "names": ["chr1", "chr2"]})
    # failures: a sequence shorter than k, line width 0, a CDS past the end of its sequence
    short = {"genome": ">a\nACGTACGTACGTAAA\n>b\nACG\n", "dump": ">9\nACGTA\n", "gff": "", "sam": "", "cds": "",
             "opts": {"width": 60, "soft": False, "alt": "n", "gene": False, "similarity": 0.8, "freq": 2,
                      "use_dump": True, "use_gff": False, "use_sam": False}, "k": 5, "names": ["a", "b"]}
    zero = dict(short, genome=">a\nACGTACGTACGTAAA\n", opts=dict(short["opts"], width=0))
    past = dict(short, genome=">a\nACGTACGTACGTAAA\n", gff="a\ts\tCDS\t40\t50\t.\t+\t.\tID=x\n",
                opts=dict(short["opts"], use_gff=True))
    for i, c in enumerate(cases + [short, zero, past]):
        d = os.path.join(root, f"case{i:02d}")
        os.makedirs(d)
        args = synth.mask_args(c, d, "expected.fa")
        for f in ("kmers.fa", "ann.gff3", "cds.sam", "cds.fa"):  # keep only the inputs the flags name
            if f not in args:
                os.remove(os.path.join(d, f))
        r = bind.reference_mask_genome(d, args)
        meta = {"args": args, "status": r.returncode, "stdout": r.stdout, "opts": c["opts"]}
        if r.returncode != 0 and os.path.exists(os.path.join(d, "expected.fa")):
            os.remove(os.path.join(d, "expected.fa"))
        with open(os.path.join(d, "case.json"), "w") as f:
            json.dump(meta, f, indent=1)
    print("mask cases:", len(cases) + 3)


def write_cut1gap_batch(here):
    """A GNU-parallel style command file of cut1Gap lines (README.md:117-122)
    over 14 region files of two masked genomes, and the files and stdout of
    the reference's cut1Gap run line by line in one folder: two option groups,
    two lines appending to one output prefix, a missing input file."""
    root = os.path.join(here, "cut1gap_batch")
    if os.path.isdir(root):
        shutil.rmtree(root)
    os.makedirs(root)
    rng = np.random.default_rng(20261205)

    def masked(codes, nruns):
        s = list(synth.to_text(codes))
        for _ in range(nruns):
            p = int(rng.integers(0, len(s) - 100))
            s[p:p + int(rng.integers(5, 120))] = "n" * len(s[p:p + int(rng.integers(5, 120))])
        for _ in range(nruns // 2):
            p = int(rng.integers(0, len(s) - 50))
            s[p:p + 25] = [c.lower() if c != "n" else c for c in s[p:p + 25]]
        return "".join(s)

    r_chr = [synth.random_seq(rng, 40000) for _ in range(2)]
    q_chr = [synth.random_seq(rng, 40000) for _ in range(2)]
    regions = []
    for t in range(14):
        rc = int(rng.integers(0, 2))
        rs = int(rng.integers(1, 40000 - 6000))
        rl = int(rng.integers(1500, 5000))
        qs = []
        for q in range(int(rng.integers(1, 5))):
            qc = int(rng.integers(0, 2))
            qs_ = int(rng.integers(1, 40000 - 6000))
            ql = int(rng.integers(1200, 5000))
            strand = int(rng.integers(0, 2))
            for _ in range(int(rng.integers(2, 8))):  # conserved fragments
                ln = int(rng.integers(40, 200))
                if ln >= min(rl, ql) - 10:
                    continue
                p1 = rs - 1 + int(rng.integers(0, rl - ln))
                p2 = qs_ - 1 + int(rng.integers(0, ql - ln))
                m = synth.mutate(rng, r_chr[rc][p1:p1 + ln], 0.08, 0.02)[:ln]
                if strand == 0:
                    m = (4 - m[::-1]).astype(np.int8)
                q_chr[qc][p2:p2 + len(m)] = m
            qs.append((f"q{q}", f"sp{chr(66 + q)}", f"qc{qc}", strand, qs_, qs_ + ql - 1))
        regions.append((("ref", "spA", f"rc{rc}", int(rng.integers(0, 2)), rs, rs + rl - 1), qs))
    fasta_with_fai(os.path.join(root, "refgenome.fa"), [(f"rc{i}", masked(c, 20)) for i, c in enumerate(r_chr)])
    fasta_with_fai(os.path.join(root, "querygenome.fa"), [(f"qc{i}", masked(c, 20)) for i, c in enumerate(q_chr)])
    lines = []
    for t, (ref, qs) in enumerate(regions):
        with open(os.path.join(root, f"region{t:02d}.fa"), "w") as fh:
            for name, sp, chrom, strand, st, en in [ref] + qs:
                fh.write(f">{name}\tspecies:{sp}\tchr:{chrom}\tstrand:{strand}\tstart:{st}\tend:{en}\nACGT\n")
        out = f"region{t:02d}.fa.5" if t != 5 else "region04.fa.5"  # line 5 appends to line 4's files
        extra = ["-c", "30"] if t >= 11 else []  # a second group
        lines.append(["dCNS", "cut1Gap", "-ra", "refgenome.fa", "-qa", "querygenome.fa", "-i", f"region{t:02d}.fa",
                      "-r", "ref", "-o", out] + extra)
    lines.insert(8, ["dCNS", "cut1Gap", "-ra", "refgenome.fa", "-qa", "querygenome.fa", "-i", "missing.fa", "-r",
                     "ref", "-o", "missing.fa.5"])
    with open(os.path.join(root, "command1"), "w") as fh:
        fh.write("".join(" ".join(l) + "\n" for l in lines))
    work = tempfile.mkdtemp()
    for f in os.listdir(root):
        shutil.copy(os.path.join(root, f), work)
    stdout, status = "", []
    for l in lines:
        r = subprocess.run([bind.REF_CLI] + l[1:], cwd=work, capture_output=True, text=True, timeout=600)
        stdout += r.stdout
        status.append(r.returncode)
    for f in sorted(os.listdir(work)):
        if f.endswith(".sam") or f.endswith("o") and ".fa.5" in f:
            shutil.copy(os.path.join(work, f), os.path.join(root, "expected_" + f))
    with open(os.path.join(root, "expected_stdout.txt"), "w") as fh:
        fh.write(stdout)
    with open(os.path.join(root, "expected_status.json"), "w") as fh:
        json.dump(status, fh)
    print("cut1gap batch: lines", len(lines), "status", status, "outputs",
          len([f for f in os.listdir(root) if f.startswith("expected_region")]))


if __name__ == "__main__":
    if "--mask" in sys.argv[1:]:
        write_mask(HERE)
    elif "--batch" in sys.argv[1:]:
        write_cut1gap_batch(HERE)
    elif "--weighted" in sys.argv[1:] or "--seeded" in sys.argv[1:]:
        r = bind.reference()
        if r is None:
            sys.exit("oracle/_ref is not built")
        if "--weighted" in sys.argv[1:]:
            write_weighted(HERE, r)
        if "--seeded" in sys.argv[1:]:
            write_seeded(HERE, r)
    else:
        main()

"""The two-piece path on the GPU: SemiGlobal (K4) tasks against the golden
vectors of the compiled reference and the oracle, findSimilarFragments
overload 4 against the reference's fragments and the oracle, and the
cut2Gaps / pairCnsXExtend / pairCns2Gaps commands.  Integer outputs are
compared exactly; p-values are the same double expression on the same
integers, so they are compared exactly too."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import dcns_amd
from dcns_amd import synth

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _params(sc2, zdrop, bandwidth, **kw):
    return dcns_amd.default_params(match=sc2[0], mismatch=sc2[1], open=sc2[2], ext=sc2[3], open2=sc2[4],
                                   ext2=sc2[5], zdrop=zdrop, bandwidth=bandwidth, **kw)


def _rc(a):
    return (4 - np.asarray(a, np.int8)[::-1]).astype(np.int8)


def test_semiglobal_golden(gpu):
    cases = json.load(open(os.path.join(GOLD, "k234_extension.json")))
    print("semiglobal golden", len(cases), flush=True)
    for c in cases:
        a, b = synth.encode(c["s1"]), synth.encode(c["s2"])
        p = _params(c["sc2"], c["zdrop"], c["bandwidth"])
        out, cig = gpu.semiglobal([(a, b)], [0], [(0, 0)], params=p)
        assert [int(v) for v in out[0]] == c["semiglobal"], c
        assert cig[0] == c["semiglobal_cigar"], c


@pytest.mark.parametrize("bandwidth,zdrop,sc2", [
    (20000, 50, (2, -3, -4, -2, -45, 0)),
    (150, 50, (2, -3, -4, -2, -45, 0)),
    (40, 25, (2, -3, -4, -2, -20, -1)),
    (3, 10, (1, -2, -3, -1, -15, 0)),
    # many equal candidates: the tie order of the five-way maximum and of
    # extension against opening decides the trace
    (60, 20, (1, -1, -2, -1, -3, 0)),
    # penalties far from the defaults, scaled by 256 in the tagged traceback
    (40, 30, (5, -4, -120, -7, -5000, -1)),
])
def test_semiglobal_tasks_vs_oracle(gpu, oracle_lib, bandwidth, zdrop, sc2):
    """Forward and reverse tasks from anchors inside batched pairs, some with
    more rows than columns + band (the rows past the band's reach)."""
    rng = np.random.default_rng(bandwidth + zdrop)
    pairs, tasks, anchors, rev = [], [], [], []
    for k in range(6):
        a = synth.random_seq(rng, int(rng.integers(200, 1400)))
        b = synth.mutate(rng, a[int(rng.integers(0, 100)):], 0.12, 0.03)
        if k % 3 == 2:
            b = b[:int(rng.integers(20, 200))]  # short query: R > C + w for the narrow bands
        pairs.append((a, b))
        for _ in range(3):
            a1, a2 = int(rng.integers(0, len(a))), int(rng.integers(0, len(b)))
            for r in (0, 1):
                tasks.append(k)
                anchors.append((a1, a2))
                rev.append(r)
    p = _params(sc2, zdrop, bandwidth)
    print("semiglobal tasks", len(tasks), bandwidth, zdrop, flush=True)
    out, cig = gpu.semiglobal(pairs, tasks, anchors, rev, params=p)
    for i, (k, (a1, a2), r) in enumerate(zip(tasks, anchors, rev)):
        a, b = pairs[k]
        s1, s2 = (_rc(a[:a1 + 1]), _rc(b[:a2 + 1])) if r else (a[a1:], b[a2:])
        want, wc = oracle_lib.semiglobal(s1, s2, sc2, zdrop, bandwidth)
        if len(s1) == 0 or len(s2) == 0:
            continue
        assert tuple(int(v) for v in out[i]) == want, (i, k, a1, a2, r)
        assert cig[i] == wc, (i, k, a1, a2, r)


@pytest.mark.parametrize("sc2", [(2, -3, -4, -2, -45, 0), (2, -3, -128, -128, -1000000, -10000)])
def test_semiglobal_wide_query_vs_oracle(gpu, oracle_lib, sc2):
    """A 2 kb reference against a 17 kb query holding a mutated copy of it,
    forward and reverse tasks from anchors along the copy.  With the default
    scoring every in-band score fits the kernel's keyed row maximum (score
    and column in one int).  With gap costs of -128 per base and a second
    piece that never helps, a gap across the query costs ~2.2e6, past the
    22-bit bound, so the launch takes the kernel that carries the maximum
    and its column apart.  Both are compared with the oracle, cigars
    included."""
    rng = np.random.default_rng(17)
    a = synth.random_seq(rng, 2000)
    b = np.concatenate([synth.random_seq(rng, 7000), synth.mutate(rng, a, 0.1, 0.0), synth.random_seq(rng, 8000)])
    # anchors on matching bases of the copy (a SemiGlobal whose first row has
    # no positive cell ends there), plus the two sequence ends
    on = [x for x in range(0, 2000, 37) if a[x] == b[7000 + x]]
    anchors = [(x, 7000 + x) for x in (on[0], on[len(on) // 3], on[2 * len(on) // 3], on[-1])]
    anchors += [(0, 0), (1999, len(b) - 1)]
    tasks = [0] * (2 * len(anchors))
    anc = [x for x in anchors for _ in (0, 1)]
    rev = [r for _ in anchors for r in (0, 1)]
    p = _params(sc2, 50, 20000)
    print("wide query", sc2, flush=True)
    out, cig = gpu.semiglobal([(a, b)], tasks, anc, rev, params=p)
    for i, ((a1, a2), r) in enumerate(zip(anc, rev)):
        s1, s2 = (_rc(a[:a1 + 1]), _rc(b[:a2 + 1])) if r else (a[a1:], b[a2:])
        want, wc = oracle_lib.semiglobal(s1, s2, sc2, 50, 20000)
        assert tuple(int(v) for v in out[i]) == want, (i, a1, a2, r)
        assert cig[i] == wc, (i, a1, a2, r)


def _cmp_frags(got, want):
    assert len(got) == len(want)
    for g, w in zip(got, want):
        assert list(g[:5]) == list(w[:5]) and g[5] == w[5] and g[6] == w[6] and list(g[7]) == list(w[7])


def test_two_piece_golden(gpu):
    cases = json.load(open(os.path.join(GOLD, "two_piece_fragments.json")))
    for c in cases:
        kw = dict(c["params"])
        extra = {"min_score": kw.pop("min_score")} if "min_score" in kw else {}
        p = _params(kw["sc"], kw["zdrop"], kw["bandwidth"], **extra)
        print("two-piece golden", kw, len(c["s1"]), len(c["s2"]), flush=True)
        res = gpu.find_similar_2gaps([(synth.encode(c["s1"]), synth.encode(c["s2"]))], params=p)
        _cmp_frags(res[0][1], c["fragments"])


def test_two_piece_batch_vs_oracle(gpu, oracle_lib):
    pairs = [synth.one_pair(seed=4400 + k, L1=int(1500 + 700 * k), L2=int(2500 - 300 * k), nfrag=5 + k)[:2]
             for k in range(5)]
    print("two-piece batch", [(len(a), len(b)) for a, b in pairs], flush=True)
    res = gpu.find_similar_2gaps(pairs)
    st = gpu.stats()
    total = 0
    for (a, b), r in zip(pairs, res):
        o = {}
        want = oracle_lib.find_similar_2gaps(a, b, stats=o)
        _cmp_frags(r[1], want)
        total += o["semiglobal_cells"]
    assert st["semiglobal_cells"] == total


def _run(args, cwd):
    print("dCNS", " ".join(args), flush=True)
    r = subprocess.run([dcns_amd.CLI_PATH] + args, cwd=cwd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return r.stdout


@pytest.mark.parametrize("cmd,args,tag", [
    ("cut2Gaps", ["-i", "in.fa", "-r", "ref", "-ra", "refgenome.fa", "-qa", "querygenome.fa"], "c2_"),
    ("cut2Gaps", ["-i", "in_refminus.fa", "-r", "ref", "-ra", "refgenome.fa", "-qa", "querygenome.fa",
                  "-O2", "-30", "-E2", "-1", "-u2", "500"], "c2m_"),
    ("pairCnsXExtend", ["-i", "pair.fa", "-r", "ref"], "px_"),
    ("pairCnsXExtend", ["-i", "pair_refminus.fa", "-r", "ref", "-O", "-3", "-E", "-1"], "pxm_"),
])
def test_commands_match_reference_files(tmp_path, cmd, args, tag):
    src = os.path.join(GOLD, "cut2gaps")
    for f in os.listdir(src):
        if not f.startswith("expected_"):
            shutil.copy(os.path.join(src, f), tmp_path)
    stdout = _run([cmd] + args + ["-o", tag], tmp_path)
    assert stdout == open(os.path.join(src, f"expected_stdout_{tag}.txt")).read()
    expected = sorted(f[len("expected_"):] for f in os.listdir(src) if f.startswith("expected_" + tag + "sp"))
    produced = sorted(f for f in os.listdir(tmp_path) if f.startswith(tag + "sp"))
    assert produced == expected and expected
    for f in expected:
        assert open(os.path.join(tmp_path, f)).read() == open(os.path.join(src, "expected_" + f)).read(), f


def test_paircns2gaps_against_oracle(tmp_path, oracle_lib):
    """pairCns2Gaps: the input's sequences through overload 4, written by the
    outputNoMasking writer that pairCnsXExtend pins against the reference."""
    src = os.path.join(GOLD, "cut2gaps")
    shutil.copy(os.path.join(src, "pair.fa"), tmp_path)
    _run(["pairCns2Gaps", "-i", "pair.fa", "-r", "ref", "-o", "p2_"], tmp_path)
    recs = open(os.path.join(src, "pair.fa")).read().split(">")[1:]
    seqs = {r.split("\t")[0]: r.split("\n")[1] for r in recs}
    ref = synth.encode(seqs["ref"])
    for name, sp in (("q0", "spB"), ("q1", "spC"), ("q2", "spD")):
        lines = open(tmp_path / f"p2_{sp}o").read().splitlines()
        assert lines[0].startswith("#cr:1:101-cq") and lines[-1] == "#done"
        want = oracle_lib.find_similar_2gaps(ref, synth.encode(seqs[name]))
        got = [tuple(int(x) for x in ln.split()[:5]) for ln in lines[1:-1]]
        assert got == [tuple(w[:5]) for w in want]
        assert len(open(tmp_path / f"p2_{sp}.sam").read().splitlines()) == len(want)


def test_xdrop_wide_bands_vs_oracle(gpu, oracle_lib):
    """Cheap gaps make the reference double the x-drop band past 161 cells;
    those tasks run in the scratch-band kernels."""
    recs = open(os.path.join(GOLD, "cut2gaps", "pair_refminus.fa")).read().split(">")[1:]
    seqs = [synth.encode(r.split("\n")[1]) for r in recs]
    pairs = [(seqs[0], q) for q in seqs[1:]] + [synth.one_pair(seed=4500 + k, L1=3000, L2=2500, nfrag=6)[:2]
                                                 for k in range(3)]
    for sc in ((2, -3, -3, -1), (1, -1, -2, -1)):
        p = dcns_amd.default_params(match=sc[0], mismatch=sc[1], open=sc[2], ext=sc[3], min_score=30)
        print("x-drop wide bands", sc, flush=True)
        res = gpu.find_similar(pairs, params=p)
        for (a, b), r in zip(pairs, res):
            want = oracle_lib.find_similar(a, b, sc=sc, min_score=30)
            _cmp_frags(r[1], want)


def test_two_piece_long_masked_vs_fixture(gpu):
    """Overload 4 on a long-masked pair of the benchmark's shape (3 kb repeat
    copies at 10% divergence, fragments 25-2,000 bp apart, n runs stripped)
    at 13-14 kb: fragments and reference cell counts equal the oracle's
    (tests/golden/two_piece_masked.json, checked against the compiled
    reference when written), and the alignments take the second gap piece:
    gap runs of 22 bp or more, which -4 / -2 alone scores below -45."""
    import importlib.util
    import json
    import os
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    spec = importlib.util.spec_from_file_location("mk", os.path.join(here, "make_two_piece_masked.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    fx = json.load(open(os.path.join(here, "two_piece_masked.json")))
    a, b = mk.pair()
    assert (len(a), len(b)) == (fx["len1"], fx["len2"])
    print("two-piece long masked pair", len(a), len(b), flush=True)
    got = gpu.find_similar_2gaps([(a, b)])[0][1]
    st = gpu.stats()
    want = fx["fragments"]
    assert len(got) == len(want)
    for g, w in zip(got, want):
        assert list(g[:5]) == w[:5] and g[5] == w[5] and g[6] == w[6] and list(g[7]) == w[7]
    for k in ("scan_cells", "xdrop_cells", "semiglobal_cells", "seeds"):
        assert st[k] == fx["stats"][k], k
    long_runs = sum(1 for f in got for c in f[7] if (c & 15) in (1, 2) and (c >> 4) >= 22)
    assert long_runs >= 10

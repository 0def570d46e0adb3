"""The weighted (ZDP) paths on the GPU: weighted SemiGlobal tasks (two-piece
and single gap) against the golden vectors of the compiled reference and the
oracle, the _wighted / _wighted_1gap overloads against the reference's
fragments and the oracle, and the weighted1Gap / weighted2Gaps commands
against the reference's files.  Integer outputs and p-values are compared
exactly."""
import json
import os
import re
import shutil
import subprocess
from collections import defaultdict

import numpy as np
import pytest

import dcns_amd
from dcns_amd import synth

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def fixtures():
    d = json.load(open(os.path.join(GOLD, "weighted.json")))
    d["prof"] = {k: {int(c): tuple(v) for c, v in p.items()} for k, p in d["profiles"].items()}
    return d


def _rc(a):
    return (4 - np.asarray(a, np.int8)[::-1]).astype(np.int8)


def _frags(res):
    return [list(g[:5]) + [g[5], g[6], list(g[7])] for g in res]


def test_weighted_tasks_golden(gpu, fixtures):
    groups = defaultdict(list)
    for t in fixtures["tasks"]:
        groups[(t["profiles"], t["w"], t["two_piece"])].append(t)
    for (name, w, tp), ts in groups.items():
        print("weighted golden tasks", name, w, tp, len(ts), flush=True)
        pairs = [(synth.encode(t["s1"]), synth.encode(t["s2"])) for t in ts]
        wc = [t["wcat"] for t in ts]
        p = dcns_amd.default_params(bandwidth=w)
        n = len(ts)
        out, cig = gpu.semiglobal_weighted(pairs, wc, fixtures["prof"][name], list(range(n)), [(0, 0)] * n,
                                           two_piece=tp, params=p)
        so, _ = gpu.semiglobal_weighted(pairs, wc, fixtures["prof"][name], list(range(n)), [(0, 0)] * n,
                                        two_piece=tp, params=p, with_cigar=False)
        for i, t in enumerate(ts):
            assert [int(v) for v in out[i]] == t["out"] and cig[i] == t["cigar"], t
            assert [int(v) for v in so[i]] == t["out_score_only"], t


@pytest.mark.parametrize("bandwidth,two_piece", [(20000, True), (20000, False), (150, True), (40, False),
                                                 (3, True)])
def test_weighted_tasks_vs_oracle(gpu, oracle_lib, fixtures, bandwidth, two_piece):
    """Forward and reverse tasks from anchors inside batched pairs; the
    reverse tasks read the categories backwards, as the reference's
    weight_rev pointer does."""
    prof = fixtures["prof"]["alt" if two_piece else "default"]
    cats = np.array(list(prof), np.int16)
    rng = np.random.default_rng(bandwidth + 7 * two_piece)
    pairs, wcs, tasks, anchors, rev = [], [], [], [], []
    for k in range(6):
        a = synth.random_seq(rng, int(rng.integers(200, 1500)))
        b = synth.mutate(rng, a[int(rng.integers(0, 100)):], 0.12, 0.03)
        if k % 3 == 2:
            b = b[:int(rng.integers(20, 200))]
        wc = np.repeat(rng.choice(cats, len(a) // 50 + 2), 50)[:len(a) + 2]
        pairs.append((a, b))
        wcs.append(wc)
        for _ in range(3):
            a1, a2 = int(rng.integers(0, len(a))), int(rng.integers(0, len(b)))
            for r in (0, 1):
                tasks.append(k)
                anchors.append((a1, a2))
                rev.append(r)
    p = dcns_amd.default_params(bandwidth=bandwidth)
    print("weighted tasks", len(tasks), bandwidth, two_piece, flush=True)
    out, cig = gpu.semiglobal_weighted(pairs, wcs, prof, tasks, anchors, rev, two_piece=two_piece, params=p)
    for i, (k, (a1, a2), r) in enumerate(zip(tasks, anchors, rev)):
        a, b = pairs[k]
        W = wcs[k]  # W[x + 1] = category of position x
        if r:
            s1, s2, wt = _rc(a[:a1 + 1]), _rc(b[:a2 + 1]), W[0:a1 + 3][::-1]
        else:
            s1, s2, wt = a[a1:], b[a2:], W[a1:len(a) + 2]
        if len(s1) == 0 or len(s2) == 0:
            continue
        mx, e1, e2, wc, _ = oracle_lib.semiglobal_weighted(s1, s2, wt, prof, w=bandwidth, two_piece=two_piece)
        assert [int(v) for v in out[i]] == [mx, e1, e2], (i, k, a1, a2, r)
        assert cig[i] == wc, (i, k, a1, a2, r)


def test_weighted_pairs_golden(gpu, fixtures):
    for c in fixtures["pairs"]:
        a, b = synth.encode(c["s1"]), synth.encode(c["s2"])
        p = dcns_amd.default_params(**c["params"])
        print("weighted golden pair", len(a), len(b), c["two_piece"], c["params"], flush=True)
        res = gpu.find_similar_weighted([(a, b)], [c["wcat"]], fixtures["prof"][c["profiles"]],
                                        two_piece=c["two_piece"], params=p)
        assert _frags(res[0][1]) == _frags(c["fragments"])


@pytest.mark.parametrize("two_piece", [True, False])
def test_weighted_batch_vs_oracle(gpu, oracle_lib, fixtures, two_piece):
    prof = fixtures["prof"]["default"]
    cats = np.array(list(prof), np.int16)
    rng = np.random.default_rng(17 + two_piece)
    pairs, wcs = [], []
    for k in range(5):
        a, b = synth.one_pair(seed=4600 + k, L1=int(1200 + 600 * k), L2=int(2600 - 300 * k), nfrag=4 + k)[:2]
        pairs.append((a, b))
        wcs.append(np.repeat(rng.choice(cats, len(a) // 120 + 2), 120)[:len(a) + 2])
    print("weighted batch", two_piece, [(len(a), len(b)) for a, b in pairs], flush=True)
    res = gpu.find_similar_weighted(pairs, wcs, prof, two_piece=two_piece)
    st = gpu.stats()
    total = 0
    for (a, b), wc, r in zip(pairs, wcs, res):
        o = {}
        want = oracle_lib.find_similar_weighted(a, b, wc, prof, two_piece=two_piece, stats=o)
        assert _frags(r[1]) == _frags(want)
        total += o["semiglobal_cells"]
    assert st["semiglobal_cells"] == total


def test_unknown_category_is_rejected(gpu):
    a, b = synth.one_pair(seed=1, L1=300, L2=300, nfrag=2)[:2]
    with pytest.raises(dcns_amd.DcnsError):
        gpu.find_similar_weighted([(a, b)], [np.full(len(a) + 2, 55)], {10: (2, -3, -4, -2, -45, 0, 50)})


_STAMP = re.compile(r"^(Mon|Tue|Wed|Thu|Fri|Sat|Sun) (Jan|Feb|Mar|Apr|May|Jun|Jul|Aug|Sep|Oct|Nov|Dec) [ \d]\d "
                    r"\d\d:\d\d:\d\d \d{4}$")


def _without_stamps(text):
    return [ln for ln in text.splitlines() if not _STAMP.match(ln)]


@pytest.mark.parametrize("cmd,args,tag", [
    ("weighted1Gap", ["-sp", "ref_scores"], "w1_"),
    ("weighted2Gaps", ["-sp", "ref_scores"], "w2_"),
    ("weighted1Gap", ["-sp", "alt_scores", "-O1", "-3", "-E1", "-1"], "w1a_"),
    ("weighted2Gaps", ["-sp", "alt_scores", "-u2", "300"], "w2a_"),
])
def test_weighted_commands_match_reference_files(tmp_path, cmd, args, tag):
    src = os.path.join(GOLD, "weighted_cmd")
    for f in os.listdir(src):
        p = os.path.join(src, f)
        if os.path.isdir(p):
            shutil.copytree(p, tmp_path / f)
        elif not f.startswith("expected_"):
            shutil.copy(p, tmp_path)
    full = [cmd, "-i", "in.fa", "-r", "ref", "-ra", "refgenome.fa", "-f", "ann.gff3"] + args + ["-o", tag]
    print("dCNS", " ".join(full), flush=True)
    r = subprocess.run([dcns_amd.CLI_PATH] + full, cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    want = open(os.path.join(src, f"expected_stdout_{tag}.txt")).read()
    assert _without_stamps(r.stdout) == _without_stamps(want)
    assert len(r.stdout.splitlines()) == len(want.splitlines())
    expected = sorted(f[len("expected_"):] for f in os.listdir(src) if f.startswith("expected_" + tag + "sp"))
    produced = sorted(f for f in os.listdir(tmp_path) if f.startswith(tag + "sp"))
    assert produced == expected and expected
    for f in expected:
        assert open(os.path.join(tmp_path, f)).read() == open(os.path.join(src, "expected_" + f)).read(), f

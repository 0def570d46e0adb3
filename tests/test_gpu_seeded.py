"""Overload 3 (the cut2Gaps2 path) on the GPU: batched seeded SemiGlobal
against the reference's fragments and the oracle, and the cut2Gaps2 command
against the reference's files."""
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


def _frags(res):
    return [list(g[:5]) + [g[5], g[6], list(g[7])] for g in res]


def _params(kw):
    sc = kw.get("sc", (2, -3, -4, -2, -45, 0))
    return dcns_amd.default_params(match=sc[0], mismatch=sc[1], open=sc[2], ext=sc[3], open2=sc[4], ext2=sc[5],
                                   zdrop=kw.get("zdrop", 50), bandwidth=kw.get("bandwidth", 20000))


def test_seeded_golden(gpu):
    cases = json.load(open(os.path.join(GOLD, "seeded_fragments.json")))
    for c in cases:
        kw = c["params"]
        print("seeded golden", len(c["s1"]), len(c["s2"]), len(c["seeds"]), kw, flush=True)
        res = gpu.find_similar_seeded([(synth.encode(c["s1"]), synth.encode(c["s2"]))], [c["seeds"]],
                                      max_align_length=kw["max_len"], params=_params(kw))
        assert _frags(res[0][1]) == _frags(c["fragments"])


@pytest.mark.parametrize("max_len", [140000, 300])
def test_seeded_batch_vs_oracle(gpu, oracle_lib, max_len):
    rng = np.random.default_rng(max_len)
    pairs, seeds = [], []
    for k in range(5):
        a, b = synth.one_pair(seed=4700 + k, L1=int(1200 + 500 * k), L2=int(2400 - 250 * k), nfrag=5 + k)[:2]
        sd = []
        for f in oracle_lib.find_similar(a, b):
            d = int(rng.integers(-2, 3))
            sd.append((f[0] - 1, min(f[1] - 1 + d, len(a) - 1), f[2] - 1, min(f[3] - 1 + d, len(b) - 1)))
        sd += [(0, 0, 0, 0), (5, len(a) + 3, 5, 10)] + sd[:1]  # a seed at the start, one outside, a duplicate
        pairs.append((a, b))
        seeds.append(sd)
    print("seeded batch", max_len, [len(s) for s in seeds], flush=True)
    res = gpu.find_similar_seeded(pairs, seeds, max_align_length=max_len)
    st = gpu.stats()
    total = 0
    for (a, b), sd, r in zip(pairs, seeds, res):
        o = {}
        want = oracle_lib.find_similar_seeded(a, b, sd, max_len=max_len, stats=o)
        assert _frags(r[1]) == _frags(want)
        total += o["semiglobal_cells"]
    assert st["semiglobal_cells"] == total


@pytest.mark.parametrize("args", [["-o", "g2_", "-m", "12000"],
                                  ["-o", "g2s_", "-m", "3000", "-u2", "500", "-O2", "-30", "-E2", "-1"]])
def test_cut2gaps2_matches_reference_files(tmp_path, args):
    src = os.path.join(GOLD, "cut2gaps2")
    for f in os.listdir(src):
        if not f.startswith("expected_"):
            shutil.copy(os.path.join(src, f), tmp_path)
    full = ["cut2Gaps2", "-sa", "in.sam", "-ra", "refgenome.fa", "-qa", "querygenome.fa"] + args
    print("dCNS", " ".join(full), flush=True)
    r = subprocess.run([dcns_amd.CLI_PATH] + full, cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    tag = args[1]
    assert r.stdout == open(os.path.join(src, f"expected_stdout_{tag}.txt")).read()
    for f in (tag + ".sam", tag + "o"):
        assert open(tmp_path / f).read() == open(os.path.join(src, "expected_" + f)).read(), f

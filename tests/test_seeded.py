"""Overload 3 (the cut2Gaps2 path) on the CPU: the oracle against the golden
fragments of the compiled reference and against the live reference on
seeds built from real fragments, moved ends (the reference's end
correction), random and duplicate seeds, and capped DP lengths."""
import json
import os

import numpy as np

from dcns_amd import synth

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_oracle_matches_golden(oracle_lib):
    for c in json.load(open(os.path.join(GOLD, "seeded_fragments.json"))):
        kw = dict(c["params"])
        if "sc" in kw:
            kw["sc"] = tuple(kw["sc"])
        got = oracle_lib.find_similar_seeded(synth.encode(c["s1"]), synth.encode(c["s2"]), c["seeds"], **kw)
        assert [list(g[:5]) + [g[5], g[6], list(g[7])] for g in got] == \
            [list(w[:5]) + [w[5], w[6], list(w[7])] for w in c["fragments"]]


def test_oracle_vs_live_reference(oracle_lib, ref_lib):
    rng = np.random.default_rng(3)
    for t in range(16):
        a, b = synth.one_pair(seed=300 + t, L1=int(rng.integers(800, 2500)), L2=int(rng.integers(800, 2500)),
                              nfrag=5)[:2]
        seeds = []
        for f in oracle_lib.find_similar(a, b):
            d = int(rng.integers(-3, 4))
            seeds.append((f[0] - 1, min(f[1] - 1 + d, len(a) - 1), f[2] - 1,
                          min(f[3] - 1 + (d if rng.random() < 0.5 else 0), len(b) - 1)))
        for _ in range(3):
            e1, e2 = int(rng.integers(0, len(a))), int(rng.integers(0, len(b)))
            seeds.append((max(0, e1 - 30), e1, max(0, e2 - 30), e2))
        seeds += [(0, 0, 0, 0), seeds[0]]
        kw = dict(max_len=int(rng.choice([140000, 500, 120])), bandwidth=int(rng.choice([20000, 60])),
                  sc=[(2, -3, -4, -2, -45, 0), (2, -3, -4, -2, -20, -1)][t % 2])
        assert oracle_lib.find_similar_seeded(a, b, seeds, **kw) == ref_lib.find_similar_seeded(a, b, seeds, **kw), t

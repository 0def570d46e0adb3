"""The GPU path directly against the golden fixtures from the compiled
reference (no oracle in between)."""
import json
import os

import numpy as np
import pytest

from dcns_amd import synth

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return json.load(open(os.path.join(GOLD, name)))


def test_k1_golden(gpu):
    cases = [c for c in load("k1_smithwaterman.json") if not c["textbook"]]
    for sc in {tuple(c["sc"]) for c in cases}:
        sub = [c for c in cases if tuple(c["sc"]) == sc]
        print("k1 golden", sc, len(sub), flush=True)
        got = gpu.sw_batch([(synth.encode(c["ref"]), synth.encode(c["read"])) for c in sub], sc)
        for g, c in zip(got, sub):
            want = c["out"]
            assert int(g[0]) == want[0]
            if want[0] > 0:
                assert [int(v) for v in g] == want


def test_pipeline_golden(gpu):
    cases = load("pipeline_fragments.json")
    pairs = [(synth.encode(c["s1"]), synth.encode(c["s2"])) for c in cases]
    print("pipeline golden", len(pairs), flush=True)
    seeds = gpu.seed_sets(pairs)
    linked = gpu.seed_sets(pairs, linked=True)
    frags = gpu.find_similar(pairs)
    for c, s, l, f in zip(cases, seeds, linked, frags):
        assert [list(x) for x in s] == c["seeds"]
        assert [list(x) for x in l] == c["linked"]
        assert f[0] == len(c["seeds"])
        assert len(f[1]) == len(c["fragments"])
        for g, w in zip(f[1], c["fragments"]):
            assert list(g[:5]) == w[:5] and g[5] == w[5] and g[6] == w[6] and g[7] == w[7]

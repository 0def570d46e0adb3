"""The CPU oracle against the golden fixtures written from the compiled
reference (tests/golden/make_golden.py).  CPU only."""
import json
import os

import numpy as np
import pytest

from dcns_amd import synth

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return json.load(open(os.path.join(GOLD, name)))


def enc(t):
    return synth.encode(t)


def test_k1_fixtures(oracle_lib):
    for case in load("k1_smithwaterman.json"):
        got = list(oracle_lib.sw(enc(case["ref"]), enc(case["read"]), tuple(case["sc"]), case["textbook"]))
        want = case["out"]
        assert got[0] == want[0], case["sc"]
        if want[0] > 0:
            assert got == want, case["sc"]


def test_xdrop_and_semiglobal_fixtures(oracle_lib):
    for case in load("k234_extension.json"):
        a, b, sc = enc(case["s1"]), enc(case["s2"]), tuple(case["sc"])
        assert list(oracle_lib.xextend(a, b, sc, case["xdrop"], case["w"])) == case["xextend"]
        tb, cig = oracle_lib.xextend_tb(a, b, sc, case["xdrop"], case["w"])
        assert list(tb) == case["xextend_tb"] and cig == case["cigar"]
        sg, sgc = oracle_lib.semiglobal(a, b, tuple(case["sc2"]), case["zdrop"], case["bandwidth"])
        assert list(sg) == case["semiglobal"] and sgc == case["semiglobal_cigar"]


def test_seed_and_fragment_fixtures(oracle_lib):
    for case in load("pipeline_fragments.json"):
        a, b = enc(case["s1"]), enc(case["s2"])
        assert [list(s) for s in oracle_lib.seed_set(a, b)] == case["seeds"]
        assert [list(s) for s in oracle_lib.seed_set(a, b, linked=True)] == case["linked"]
        got = oracle_lib.find_similar(a, b)
        want = case["fragments"]
        assert len(got) == len(want)
        for g, w in zip(got, want):
            assert list(g[:5]) == w[:5] and g[5] == w[5] and g[6] == w[6] and g[7] == w[7]


def test_two_piece_fixtures(oracle_lib):
    """findSimilarFragmentsForPairedSequence overload 4 (cut2Gaps path)."""
    for case in load("two_piece_fragments.json"):
        kw = dict(case["params"])
        kw["sc"] = tuple(kw["sc"])
        got = oracle_lib.find_similar_2gaps(enc(case["s1"]), enc(case["s2"]), **kw)
        assert len(got) == len(case["fragments"])
        for g, w in zip(got, case["fragments"]):
            assert list(g[:5]) == w[:5] and g[5] == w[5] and g[6] == w[6] and list(g[7]) == w[7]

"""The benchmark's 4,096-pair batch (lengths log-uniform 1-50 kb, GC
0.40-0.60, 0-30 planted fragments) checked at its real sizes: a 96-pair
slice with the longest and the most fragment-dense pairs, through one
batched call and through a staged run of the whole batch (the path the
benchmark times), against the oracle's fragments, seed counts and reference
cell counts (tests/golden/batch_slice.json, each pair also checked against
the compiled reference when it was written)."""
import json
import os

import numpy as np
import pytest

from dcns_amd import synth

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "batch_slice.json")


@pytest.fixture(scope="module")
def batch_and_fixture():
    fx = json.load(open(GOLD))
    print("generating the 4096-pair batch", flush=True)
    pairs = synth.batch()
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_batch_slice", os.path.join(os.path.dirname(GOLD),
                                                                                    "make_batch_slice.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    assert mk.input_digest(pairs, fx["indices"]) == fx["input_sha256"], "batch generator changed"
    return pairs, fx


def _same(got, want):
    assert got[0] == want["seeds_size"], (want["index"], got[0], want["seeds_size"])
    assert len(got[1]) == len(want["fragments"]), (want["index"], len(got[1]), len(want["fragments"]))
    for g, w in zip(got[1], want["fragments"]):
        assert list(g[:5]) == w[:5], (want["index"], g[:5], w[:5])
        assert g[5] == w[5] and g[6] == w[6], (want["index"], g[5:7], w[5:7])
        assert list(g[7]) == w[7], want["index"]


def test_slice_shape(batch_and_fixture):
    pairs, fx = batch_and_fixture
    L = [max(p["len1"], p["len2"]) for p in fx["pairs"]]
    assert len(fx["pairs"]) == 96 and max(L) >= 49000
    assert sum(min(p["len1"], p["len2"]) >= 40000 for p in fx["pairs"]) >= 2
    assert all(p["reference_checked"] for p in fx["pairs"])
    assert sum(len(p["fragments"]) for p in fx["pairs"]) > 500


def test_batch_slice_one_call(gpu, batch_and_fixture):
    pairs, fx = batch_and_fixture
    sub = [pairs[i] for i in fx["indices"]]
    print("find_similar on the 96-pair slice", flush=True)
    got = gpu.find_similar(sub)
    st = gpu.stats()
    for g, w in zip(got, fx["pairs"]):
        _same(g, w)
    assert st["scan_cells"] == sum(p["scan_cells"] for p in fx["pairs"])
    assert st["xdrop_cells"] == sum(p["xdrop_cells"] for p in fx["pairs"])
    assert st["seeds"] == sum(p["seeds_size"] for p in fx["pairs"])
    assert st["linked_seeds"] == sum(p["linked"] for p in fx["pairs"])


def test_whole_batch_staged_run(gpu, batch_and_fixture):
    """The staged run of all 4,096 pairs (the benchmark's timed path: pools
    grown to the batch, every pair in one launch per level), twice, its slice
    pairs against the fixture."""
    pairs, fx = batch_and_fixture
    print("staged run of the whole batch", flush=True)
    gpu.stage(pairs)
    first = gpu.find_similar_staged()
    second = gpu.find_similar_staged()
    for i, w in zip(fx["indices"], fx["pairs"]):
        _same(first[i], w)
    assert first == second
    assert sum(len(f) for _, f in first) > 30000

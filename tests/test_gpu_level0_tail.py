"""The seed recursion with level-0 tails (level0_tail = 1): level 0 records
every row where a window's maximum reaches min_score, and a rescan runs only
its first D rows (D = wm + ceil(window * match / g)), taking the rest of its
trajectory from those rows; and level 0 in row segments (level0_split = 1),
each segment started D rows early from zeros, the segments' trajectories
merged.  Seed sets, fragments and reference cell counts must equal the
oracle's exactly, as without them; the GPU's own cell count must drop where
rescans are long."""
import numpy as np
import pytest

import dcns_amd
from dcns_amd import synth
from test_gpu_pipeline import _cmp_fragments, _small_pairs, _repeat_pair

pytestmark = pytest.mark.gpu


def _with_tail(gpu, fn, **tuning):
    tuning.setdefault("level0_split", -1)
    old = gpu.set_tuning(level0_tail=1, **tuning)
    try:
        return fn()
    finally:
        gpu.set_tuning(**old)


@pytest.mark.parametrize("split", [-1, 1])
def test_tail_masked_repeat_pair_vs_oracle(gpu, oracle_lib, split):
    t1, t2 = synth.long_masked(seed=77, L=20000, te_copies=20, te_len=1500, nfrag=30)
    a = synth.encode(t1.replace("n", ""))
    b = synth.encode(t2.replace("n", ""))
    print("tail: masked repeat pair", len(a), len(b), "split", split, flush=True)
    got = _with_tail(gpu, lambda: gpu.find_similar([(a, b)])[0], level0_split=split)
    st = gpu.stats()
    o = {}
    want = oracle_lib.find_similar(a, b, stats=o)
    assert got[0] == o["seeds"] and st["linked_seeds"] == o["linked"]
    assert st["scan_cells"] == o["scan_cells"] and st["xdrop_cells"] == o["xdrop_cells"]
    assert st["levels"] >= 3
    _cmp_fragments(got[1], want, "tail masked")
    old = gpu.set_tuning(level0_tail=-1)
    try:
        gpu.find_similar([(a, b)])
        full = gpu.stats()["gpu_scan_cells"]
    finally:
        gpu.set_tuning(**old)
    assert st["gpu_scan_cells"] < full


@pytest.mark.parametrize("split", [-1, 1])
def test_tail_long_rescans_with_n_vs_oracle(gpu, oracle_lib, split):
    rng = np.random.default_rng(4242)
    a, b = synth.one_pair(seed=4243, L1=14000, L2=6013, nfrag=16)[:2]
    for _ in range(25):
        p = int(rng.integers(0, len(a) - 60))
        a[p:p + int(rng.integers(1, 50))] = 2
    for k in range(6):
        src = int(rng.integers(0, len(b) - 300))
        dst = 1000 + k * 2100
        a[dst:dst + 300] = synth.mutate(rng, b[src:src + 300], 0.05, 0.0)[:300]
    print("tail: long rescans with N, split", split, flush=True)
    got = _with_tail(gpu, lambda: gpu.find_similar([(a, b)])[0], level0_split=split)
    st = gpu.stats()
    o = {}
    want = oracle_lib.find_similar(a, b, stats=o)
    assert got[0] == o["seeds"] and st["linked_seeds"] == o["linked"]
    assert st["scan_cells"] == o["scan_cells"] and st["xdrop_cells"] == o["xdrop_cells"]
    _cmp_fragments(got[1], want, "tail long rescans")


@pytest.mark.parametrize("sc,min_score", [((2, -3, -4, -2), 40), ((1, -1, -2, -1), 20), ((2, -3, -1, -3), 40),
                                          ((3, -4, -6, -3), 50), ((60, -80, -100, -60), 600)])
def test_tail_seed_sets_vs_oracle(gpu, oracle_lib, sc, min_score):
    """Seed sets of small pairs, tandem repeats and a repeat-rich pair under
    several scorings (gap opening cheaper and dearer than an extension); the
    last scoring's level-0 scan is not packed f16, so the tails stay off."""
    rng = np.random.default_rng(31)
    pairs = _small_pairs(91, 8, 300, 2500, (2, 12)) + [_repeat_pair(rng, 3000, 900)]
    unit = synth.random_seq(rng, 60)
    rep = np.concatenate([synth.mutate(rng, unit, 0.05, 0.0)[:60] for _ in range(40)])
    pairs += [(rep, synth.mutate(rng, unit, 0.02, 0.0)), (rep, rep[:500].copy())]
    p = dcns_amd.default_params(match=sc[0], mismatch=sc[1], open=sc[2], ext=sc[3], min_score=min_score)
    print("tail: seed sets", sc, flush=True)
    for linked, split in ((False, -1), (True, -1), (False, 1)):
        got = _with_tail(gpu, lambda: gpu.seed_sets(pairs, params=p, linked=linked), level0_split=split)
        for i, (a, b) in enumerate(pairs):
            assert got[i] == oracle_lib.seed_set(a, b, sc, min_score=min_score, linked=linked), (sc, i, linked, split)


@pytest.mark.parametrize("window,step,min_score", [(24, 6, 30), (50, 8, 60)])
def test_tail_other_windows_vs_oracle(gpu, oracle_lib, window, step, min_score):
    """Windows narrower than the 38 columns the kernels evaluate (the rest N
    padding, which the tail bound counts) and wider ones (64 columns)."""
    rng = np.random.default_rng(window)
    pairs = _small_pairs(93, 6, 400, 2500, (2, 10)) + [_repeat_pair(rng, 2500, 800)]
    p = dcns_amd.default_params(window=window, step=step, min_score=min_score)
    print("tail: window", window, flush=True)
    for split in (-1, 1):
        got = _with_tail(gpu, lambda: gpu.seed_sets(pairs, params=p, linked=False), level0_split=split)
        for i, (a, b) in enumerate(pairs):
            assert got[i] == oracle_lib.seed_set(a, b, window=window, step=step, min_score=min_score), (window, i)


@pytest.mark.parametrize("pool_min", [1, 64])
def test_tail_pools_grow_vs_oracle(gpu, oracle_lib, pool_min):
    """Level-0 tails with recursion pools that overflow on every early level
    (the resolve retries reuse the tails' trajectories)."""
    rng = np.random.default_rng(707)
    a, b = synth.one_pair(seed=708, L1=9000, L2=4000, nfrag=12)[:2]
    for k in range(5):
        src = int(rng.integers(0, len(b) - 300))
        a[500 + k * 1700:800 + k * 1700] = synth.mutate(rng, b[src:src + 300], 0.04, 0.0)[:300]
    pairs = [(a, b), _repeat_pair(rng, 4000, 1200)] + _small_pairs(71, 4)
    print("tail: pools from", pool_min, flush=True)
    got = _with_tail(gpu, lambda: gpu.find_similar(pairs), pool_min=pool_min, level0_split=1)
    st = gpu.stats()
    cells = 0
    for i, (x, y) in enumerate(pairs):
        o = {}
        want = oracle_lib.find_similar(x, y, stats=o)
        assert got[i][0] == o["seeds"], i
        _cmp_fragments(got[i][1], want, i)
        cells += o["scan_cells"]
    assert st["scan_cells"] == cells


def test_tail_level_graph_replays(gpu, oracle_lib):
    """Captured recursion levels (with the tail launch in each) replayed
    several times against the host-driven run, level-0 scan per task and as
    persistent waves."""
    rng = np.random.default_rng(919)
    pairs = _small_pairs(83, 6, 600, 5000, (1, 4)) + [_repeat_pair(rng, 6000, 1500)]
    gpu.stage(pairs)
    old = gpu.set_tuning(level0_tail=1, level0_split=1, graph_levels=-1)
    try:
        want = gpu.find_similar_staged()
        st0 = gpu.stats()
        for gl, mode, waves in [(4, 1, 0), (2, 2, 4)]:
            print("tail: level graph", gl, mode, flush=True)
            for rep in range(3):
                gpu.set_tuning(graph_levels=gl, level0_mode=mode, level0_waves=waves)
                got = gpu.find_similar_staged()
                st = gpu.stats()
                assert got == want, (gl, mode, rep)
                for k in ("scan_cells", "gpu_scan_cells", "xdrop_cells", "seeds", "linked_seeds", "levels"):
                    assert st[k] == st0[k], (gl, mode, rep, k)
                gpu.set_tuning(graph_levels=-1)
                assert gpu.find_similar_staged() == want
    finally:
        gpu.set_tuning(**old)
    for i, (x, y) in enumerate(pairs):
        o = {}
        w = oracle_lib.find_similar(x, y, stats=o)
        assert want[i][0] == o["seeds"], i
        _cmp_fragments(want[i][1], w, i)


def test_split_level0_only_vs_oracle(gpu, oracle_lib):
    """Level 0 in row segments without the tails (rescans run in full): the
    merged level-0 trajectories and lastrow feed the same recursion."""
    pairs = _small_pairs(97, 8, 1500, 9000, (2, 12))
    old = gpu.set_tuning(level0_tail=-1, level0_split=1)
    try:
        print("split only", flush=True)
        got = gpu.find_similar(pairs)
        st = gpu.stats()
    finally:
        gpu.set_tuning(**old)
    cells = 0
    for i, (x, y) in enumerate(pairs):
        o = {}
        want = oracle_lib.find_similar(x, y, stats=o)
        assert got[i][0] == o["seeds"], i
        _cmp_fragments(got[i][1], want, i)
        cells += o["scan_cells"]
    assert st["scan_cells"] == cells

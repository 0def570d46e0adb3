"""Seed sets, x-drop extension and the whole findSimilarFragmentsForPairedSequence
path on the GPU against the oracle (which is pinned to the compiled reference)."""
import numpy as np
import pytest

import dcns_amd
from dcns_amd import synth

pytestmark = pytest.mark.gpu


def _small_pairs(seed, n, lo=300, hi=3000, nfrag=(0, 12)):
    rng = np.random.default_rng(seed)
    out = []
    for t in range(n):
        L1 = int(rng.integers(lo, hi))
        L2 = int(rng.integers(lo // 3, hi))
        out.append(synth.one_pair(seed=seed * 1000 + t, L1=L1, L2=L2, nfrag=int(rng.integers(*nfrag))))
    return out


@pytest.mark.parametrize("linked", [False, True])
def test_seed_sets_match_oracle(gpu, oracle_lib, linked):
    pairs = _small_pairs(11, 24)
    print("seed_sets linked=%d" % linked, flush=True)
    got = gpu.seed_sets(pairs, linked=linked)
    for i, (a, b) in enumerate(pairs):
        assert got[i] == oracle_lib.seed_set(a, b, linked=linked), i


def test_seed_sets_repeats_and_edges(gpu, oracle_lib):
    """Tandem repeats (deep recursion), short queries, partial windows, N runs."""
    rng = np.random.default_rng(5)
    unit = synth.random_seq(rng, 60)
    rep = np.concatenate([synth.mutate(rng, unit, 0.05, 0.0)[:60] for _ in range(40)])
    q = synth.mutate(rng, unit, 0.02, 0.0)
    withn = synth.random_seq(rng, 900)
    withn[100:160] = 2
    pairs = [(rep, q), (rep, rep[:500].copy()), (synth.random_seq(rng, 500), synth.random_seq(rng, 7)),
             (withn, withn[50:400].copy()), (rep[:30].copy(), rep[:45].copy()), (q, q.copy())]
    print("seed_sets edges", flush=True)
    for linked in (False, True):
        got = gpu.seed_sets(pairs, linked=linked)
        for i, (a, b) in enumerate(pairs):
            assert got[i] == oracle_lib.seed_set(a, b, linked=linked), (i, linked)


def _x_extend_oracle(o, s1, s2, start1, start2, sc=(2, -3, -4, -2), xdrop=20, w=10):
    """x_extend_seed (tb) composed from the oracle's two SemiGlobal_xextend calls."""
    f = o.xextend(s1[start1:], s2[start2:], sc, xdrop, w)
    if f[1] < 0 or f[2] < 0:
        return (f[0], start1, 0, start2, 0, 0, None)
    r1 = (4 - s1[:start1 + f[1] + 1][::-1]).astype(np.int8)
    r2 = (4 - s2[:start2 + f[2] + 1][::-1]).astype(np.int8)
    (mx, e1, e2, iii, jjj), cig = o.xextend_tb(r1, r2, sc, xdrop, w)
    if e1 < 0 or e2 < 0:
        return (mx, start1, 0, start2, 0, 0, cig)
    end1 = start1 + f[1] - iii
    end2 = start2 + f[2] - jjj
    return (mx, start1 + f[1] - e1, end1, start2 + f[2] - e2, end2, 1, cig)


def test_xdrop_extend_matches_oracle(gpu, oracle_lib):
    pairs = _small_pairs(21, 6, 500, 2500, (5, 15))
    rng = np.random.default_rng(9)
    pos, starts = [], []
    for i, (a, b) in enumerate(pairs):
        for s in oracle_lib.seed_set(a, b, linked=True)[:40]:
            pos.append(i)
            starts.append((s[0] + 1, s[2] + 1))
        for _ in range(20):  # arbitrary starts, including near the ends
            pos.append(i)
            starts.append((int(rng.integers(0, len(a) + 1)), int(rng.integers(0, len(b) + 1))))
    print("xdrop", len(pos), flush=True)
    got = gpu.xdrop_extend(pairs, pos, starts)
    for k, (i, (s1, s2)) in enumerate(zip(pos, starts)):
        a, b = pairs[i]
        want = _x_extend_oracle(oracle_lib, a, b, s1, s2)
        g = got[k]
        if want[5] == 0:
            assert g[5] == 0, (k, g[:6], want[:6])
            continue
        assert g[:6] == want[:6], (k, g[:6], want[:6])
        assert g[6] == want[6], k


@pytest.mark.parametrize("sc,w", [((2, -3, -1, -3), 10), ((2, -3, -4, -2), 20), ((1, -1, -1, -2), 40),
                                  ((2, -3, -3, -1), 80)])
def test_xdrop_extend_wave_bands_vs_oracle(gpu, oracle_lib, sc, w):
    """Initial bands of 21-161 cells (past 21 every pass runs one wave per
    seed) and cheap gaps, open above extend as well as below, so that bands
    restart and double inside the wave kernels and the traceback pass runs
    there too."""
    pairs = _small_pairs(77, 4, 800, 3000, (6, 14))
    rng = np.random.default_rng(w)
    pos, starts = [], []
    for i, (a, b) in enumerate(pairs):
        for s in oracle_lib.seed_set(a, b, linked=True)[:30]:
            pos.append(i)
            starts.append((s[0] + 1, s[2] + 1))
        for _ in range(10):
            pos.append(i)
            starts.append((int(rng.integers(0, len(a) + 1)), int(rng.integers(0, len(b) + 1))))
    p = dcns_amd.default_params(match=sc[0], mismatch=sc[1], open=sc[2], ext=sc[3], w=w)
    print("xdrop wave bands", sc, w, len(pos), flush=True)
    got = gpu.xdrop_extend(pairs, pos, starts, params=p)
    traced = 0
    for k, (i, (s1, s2)) in enumerate(zip(pos, starts)):
        a, b = pairs[i]
        want = _x_extend_oracle(oracle_lib, a, b, s1, s2, sc=sc, w=w)
        g = got[k]
        if want[5] == 0:
            assert g[5] == 0, (k, g[:6], want[:6])
            continue
        assert g[:6] == want[:6], (k, g[:6], want[:6])
        assert g[6] == want[6], k
        traced += 1
    assert traced > 20


@pytest.mark.parametrize("w", [10, 15])
def test_xdrop_lane_row_budget_vs_oracle(gpu, oracle_lib, w):
    """The lane kernels' row budget (a pass past it moves to the wave kernel,
    a traceback pass of more rows runs there): budgets of 1, 5 and 64 rows
    and none give the oracle's results and the same reference cell count.
    w = 15 starts at a 31-cell band, on the 41-cell lane kernel."""
    pairs = _small_pairs(123, 4, 600, 2500, (6, 14))
    rng = np.random.default_rng(w + 1)
    pos, starts = [], []
    for i, (a, b) in enumerate(pairs):
        for s in oracle_lib.seed_set(a, b, linked=True)[:30]:
            pos.append(i)
            starts.append((s[0] + 1, s[2] + 1))
        for _ in range(10):
            pos.append(i)
            starts.append((int(rng.integers(0, len(a) + 1)), int(rng.integers(0, len(b) + 1))))
    p = dcns_amd.default_params(w=w)
    want = [_x_extend_oracle(oracle_lib, pairs[i][0], pairs[i][1], s1, s2, w=w) for i, (s1, s2) in zip(pos, starts)]
    old = gpu.get_tuning()
    cells = set()
    try:
        for rows in (1, 5, 64, -1):
            gpu.set_tuning(xdrop_lane_rows=rows)
            print("xdrop lane row budget", w, rows, len(pos), flush=True)
            got = gpu.xdrop_extend(pairs, pos, starts, params=p)
            cells.add(gpu.stats()["xdrop_cells"])
            for k, (g, wt) in enumerate(zip(got, want)):
                if wt[5] == 0:
                    assert g[5] == 0, (rows, k, g[:6], wt[:6])
                    continue
                assert g[:6] == wt[:6], (rows, k, g[:6], wt[:6])
                assert g[6] == wt[6], (rows, k)
    finally:
        gpu.set_tuning(**old)
    assert len(cells) == 1, cells


def _cmp_fragments(got, want, where):
    assert len(got) == len(want), (where, len(got), len(want))
    for g, w in zip(got, want):
        assert g[:5] == w[:5], (where, g[:5], w[:5])
        assert g[5] == w[5] and g[6] == w[6], (where, g[5:7], w[5:7])
        assert g[7] == w[7], where


def test_find_similar_matches_oracle(gpu, oracle_lib):
    pairs = _small_pairs(31, 16, 400, 4000, (0, 20))
    print("find_similar small", flush=True)
    got = gpu.find_similar(pairs)
    for i, (a, b) in enumerate(pairs):
        st = {}
        want = oracle_lib.find_similar(a, b, stats=st)
        assert got[i][0] == st["seeds"], i
        _cmp_fragments(got[i][1], want, i)


def test_find_similar_one_pair_10kb(gpu, oracle_lib):
    """The plain-method workload: 10 kb x 10 kb, 20 planted fragments."""
    a, b = synth.one_pair()
    print("find_similar 10kb", flush=True)
    got = gpu.find_similar([(a, b)])
    st = {}
    want = oracle_lib.find_similar(a, b, stats=st)
    assert got[0][0] == st["seeds"]
    _cmp_fragments(got[0][1], want, "10kb")
    s = gpu.stats()
    assert s["scan_cells"] == st["scan_cells"]
    assert s["xdrop_cells"] == st["xdrop_cells"]


def test_find_similar_long_rescans_with_n_vs_oracle(gpu, oracle_lib):
    """Rescans of more than 4,096 rows (one wave per task, columns across the
    lanes) on a reference with N runs left in place, a query whose last
    window is partial (the padded columns of the two-strip lane kernel), and
    planted fragments repeated along the reference so the recursion goes
    several levels deep; fragments, seed counts and cell counts equal the
    oracle's."""
    rng = np.random.default_rng(4242)
    a, b = synth.one_pair(seed=4243, L1=14000, L2=6013, nfrag=16)[:2]
    for _ in range(25):  # N runs in the reference
        p = int(rng.integers(0, len(a) - 60))
        a[p:p + int(rng.integers(1, 50))] = 2
    for k in range(6):  # copies of a query stretch along the reference: deep recursion
        src = int(rng.integers(0, len(b) - 300))
        dst = 1000 + k * 2100
        a[dst:dst + 300] = synth.mutate(rng, b[src:src + 300], 0.05, 0.0)[:300]
    assert (len(b) - 38) % 8 != 0
    print("long rescans with N", len(a), len(b), flush=True)
    got = gpu.find_similar([(a, b)])[0]
    st = gpu.stats()
    o = {}
    want = oracle_lib.find_similar(a, b, stats=o)
    assert got[0] == o["seeds"] and st["linked_seeds"] == o["linked"]
    assert st["scan_cells"] == o["scan_cells"] and st["xdrop_cells"] == o["xdrop_cells"]
    assert st["levels"] >= 3
    _cmp_fragments(got[1], want, "long rescans")


@pytest.mark.parametrize("sc,min_score", [((3, -4, -6, -3), 50), ((1, -1, -2, -1), 20), ((2, -1, -2, -1), 40),
                                          ((60, -80, -100, -60), 600), ((2, -3, -1, -3), 40),
                                          ((9, -11, -14, -5), 180)])
def test_seed_sets_other_scorings(gpu, oracle_lib, sc, min_score):
    """Scorings that take the packed-f16 window scan and (window * match >
    2048) the int16 one, including gap open cheaper than gap extend.  The
    later levels run packed too, except where the f16 patterns of match or
    mismatch have a nonzero low byte (9 = 0x4880) or the f16 bound fails:
    the int32 kernel."""
    pairs = _small_pairs(51, 10, 300, 2500, (2, 12))
    p = dcns_amd.default_params(match=sc[0], mismatch=sc[1], open=sc[2], ext=sc[3], min_score=min_score)
    print("seed_sets", sc, min_score, flush=True)
    for linked in (False, True):
        got = gpu.seed_sets(pairs, params=p, linked=linked)
        for i, (a, b) in enumerate(pairs):
            assert got[i] == oracle_lib.seed_set(a, b, sc, min_score=min_score, linked=linked), (i, linked)


def test_long_masked_repeat_pair_vs_oracle(gpu, oracle_lib):
    """The shape of the 500 kb masked pair at a size the oracle finishes in
    seconds: lowercase n runs stripped before alignment, 20 copies of a
    1.5 kb repeat at 10% divergence (deep getSeeds recursion inside every
    copy), 30 planted fragments.  Fragments, seed counts and the reference
    cell count must all match."""
    t1, t2 = synth.long_masked(seed=77, L=20000, te_copies=20, te_len=1500, nfrag=30)
    a = synth.encode(t1.replace("n", ""))
    b = synth.encode(t2.replace("n", ""))
    print("masked repeat pair", len(a), len(b), flush=True)
    got = gpu.find_similar([(a, b)])[0]
    st = gpu.stats()
    o = {}
    want = oracle_lib.find_similar(a, b, stats=o)
    assert got[0] == o["seeds"]
    assert st["seeds"] == o["seeds"] and st["linked_seeds"] == o["linked"]
    assert st["scan_cells"] == o["scan_cells"] and st["xdrop_cells"] == o["xdrop_cells"]
    assert len(got[1]) == len(want) > 50
    for g, w in zip(got[1], want):
        assert list(g[:5]) == list(w[:5]) and g[5] == w[5] and g[6] == w[6] and list(g[7]) == list(w[7])


@pytest.mark.parametrize("pool_min", [1, 64])
def test_recursion_pools_grow_vs_oracle(gpu, oracle_lib, pool_min):
    """Recursion pools that start nearly empty: every resolve launch of the
    first levels overflows its node, trajectory or task pool and is run again
    with the pools grown, on pairs whose recursion goes deep (repeats along
    the reference, a tandem repeat); seed sets, fragments and cell counts
    equal the oracle's, and the same context then runs with default pools."""
    rng = np.random.default_rng(606)
    a, b = synth.one_pair(seed=607, L1=9000, L2=4000, nfrag=12)[:2]
    for k in range(5):  # copies of a query stretch along the reference
        src = int(rng.integers(0, len(b) - 300))
        a[500 + k * 1700:800 + k * 1700] = synth.mutate(rng, b[src:src + 300], 0.04, 0.0)[:300]
    unit = synth.random_seq(rng, 70)
    rep = np.concatenate([synth.mutate(rng, unit, 0.05, 0.0)[:70] for _ in range(50)])
    pairs = [(a, b), (rep, synth.mutate(rng, unit, 0.02, 0.0)), (rep, rep[:900].copy())] + _small_pairs(61, 6)
    old = gpu.set_tuning(pool_min=pool_min)
    try:
        print("recursion pools from", pool_min, flush=True)
        got = gpu.find_similar(pairs)
        st = gpu.stats()
        sets = gpu.seed_sets(pairs, linked=False)
    finally:
        gpu.set_tuning(**old)
    cells = xcells = 0
    for i, (x, y) in enumerate(pairs):
        o = {}
        want = oracle_lib.find_similar(x, y, stats=o)
        assert got[i][0] == o["seeds"], i
        _cmp_fragments(got[i][1], want, i)
        assert sets[i] == oracle_lib.seed_set(x, y, linked=False), i
        cells += o["scan_cells"]
        xcells += o["xdrop_cells"]
    assert st["scan_cells"] == cells and st["xdrop_cells"] == xcells
    assert st["levels"] >= 4
    again = gpu.find_similar(pairs)
    assert again == got


def _repeat_pair(rng, L1, L2):
    """A reference of tandem copies of a 70 bp unit and a query holding a few
    of them: every query window over a copy recurses through all the
    reference's copies."""
    unit = synth.random_seq(rng, 70)
    ref = np.concatenate([synth.mutate(rng, unit, 0.04, 0.0)[:70] for _ in range(L1 // 70 + 1)])[:L1]
    qry = synth.random_seq(rng, L2)
    for p in range(100, L2 - 300, 700):
        qry[p:p + 210] = np.concatenate([unit, unit, unit])
    return ref.astype(np.int8), qry.astype(np.int8)


def test_level_graph_replays_vs_host_driven(gpu, oracle_lib):
    """The recursion levels replayed from a captured HIP graph (after one
    host-driven run of a new layout), with the level-0 scan per task and as
    persistent waves, with fewer captured levels than the recursion has,
    against the host-driven run: identical fragments, seeds and cell counts
    on every replay, with other device work between the replays.  Then a
    batch of the same layout whose recursion outgrows the pools the graph
    was captured with: the overflow inside the graph falls back to the
    host-driven levels, and the results equal the oracle's."""
    rng = np.random.default_rng(515)
    a_pairs = _small_pairs(81, 10, 600, 5000, (1, 4))
    a_pairs.append(synth.one_pair(seed=82, L1=9000, L2=3000, nfrag=6))
    gpu.stage(a_pairs)
    old = gpu.set_tuning(graph_levels=-1)
    try:
        want = gpu.find_similar_staged()
        st0 = gpu.stats()
        for gl, mode, waves in [(4, 1, 0), (1, 2, 4), (8, 2, 8), (2, 1, 0)]:
            print("level graph", gl, "level-0 mode", mode, flush=True)
            for rep in range(3):
                gpu.set_tuning(graph_levels=gl, level0_mode=mode, level0_waves=waves)
                got = gpu.find_similar_staged()
                st = gpu.stats()
                assert got == want, (gl, mode, rep)
                for k in ("scan_cells", "gpu_scan_cells", "xdrop_cells", "seeds", "linked_seeds", "levels"):
                    assert st[k] == st0[k], (gl, mode, rep, k)
                gpu.set_tuning(graph_levels=-1)  # eager work between the replays
                assert gpu.find_similar_staged() == want
        # pools sized for a shallow recursion, then a deep one of the same layout
        gpu.set_tuning(graph_levels=4, level0_mode=1, pool_min=64)
        for rep in range(3):
            assert gpu.find_similar_staged() == want
        b_pairs = [_repeat_pair(rng, len(x), len(y)) if k % 3 == 0 else (x, y)
                   for k, (x, y) in enumerate(a_pairs)]
        gpu.stage(b_pairs)
        print("deep recursion on the captured layout", flush=True)
        got_b = gpu.find_similar_staged()
        st_b = gpu.stats()
        again = gpu.find_similar_staged()
    finally:
        gpu.set_tuning(**old)
    assert again == got_b
    cells = 0
    for i, (x, y) in enumerate(b_pairs):
        o = {}
        w = oracle_lib.find_similar(x, y, stats=o)
        assert got_b[i][0] == o["seeds"], i
        _cmp_fragments(got_b[i][1], w, i)
        cells += o["scan_cells"]
    assert st_b["scan_cells"] == cells and st_b["levels"] >= 5

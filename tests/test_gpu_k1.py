"""K1 (SmithWaterman max + first position) on the GPU against the oracle."""
import numpy as np
import pytest

from dcns_amd import synth

pytestmark = pytest.mark.gpu

SCORINGS = [(2, -3, -4, -2), (1, -1, -2, -1), (3, -4, -6, -3), (2, -1, -2, -1), (1, -4, -2, -3)]


def _pairs(rng, n, lo, hi, related=0.5):
    out = []
    for _ in range(n):
        L1 = int(rng.integers(lo, hi + 1))
        L2 = int(rng.integers(lo, hi + 1))
        a = rng.choice([0, 1, 2, 3, 4], L1, p=[.24, .24, .04, .24, .24]).astype(np.int8)
        b = rng.choice([0, 1, 2, 3, 4], L2, p=[.24, .24, .04, .24, .24]).astype(np.int8)
        if L1 > 20 and L2 > 10 and rng.random() < related:
            k = int(rng.integers(0, L1 - 10))
            m = synth.mutate(rng, a[k:k + L2], 0.1, 0.04)[:L2]
            b[:len(m)] = m
        out.append((a, b))
    return out


@pytest.mark.parametrize("sc", SCORINGS)
def test_sw_batch_matches_oracle(gpu, oracle_lib, sc):
    rng = np.random.default_rng(abs(hash(sc)) % 2**32)
    pairs = _pairs(rng, 600, 0, 260)
    print("sw_batch", sc, len(pairs), flush=True)
    got = gpu.sw_batch(pairs, sc)
    for i, (a, b) in enumerate(pairs):
        want = oracle_lib.sw(a, b, sc)
        g = tuple(int(x) for x in got[i])
        assert g[0] == want[0], (i, len(a), len(b), g, want)
        if want[0] > 0:
            assert g == want, (i, len(a), len(b), g, want)


def test_sw_batch_long_and_textbook(gpu, oracle_lib):
    """Long pairs, including ones past the int16 bound that take the textbook recurrence."""
    rng = np.random.default_rng(7)
    a = synth.random_seq(rng, 12000)
    b = synth.mutate(rng, a, 0.05, 0.01)[:11500]
    c = synth.random_seq(rng, 3000)
    pairs = [(a, b), (a[:700], b[:2000]), (c, c.copy()), (b, a)]
    for sc in [(2, -3, -4, -2), (1, -2, -5, -3)]:
        print("sw_batch long", sc, flush=True)
        got = gpu.sw_batch(pairs, sc)
        for i, (x, y) in enumerate(pairs):
            assert tuple(int(v) for v in got[i]) == oracle_lib.sw(x, y, sc), (i, sc)


def test_identical_pair_closed_form(gpu):
    rng = np.random.default_rng(3)
    a = synth.random_seq(rng, 1500)
    got = gpu.sw_batch([(a, a.copy())])
    assert tuple(int(v) for v in got[0]) == (2 * 1500, 1499, 1499)


def _mix64(z):
    z = z + np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def _random_fragment(key, L, gc):
    """The device generator restated: base b = thresholds of 32 bits of
    mix64(key + b // 2) (low half for even b)."""
    x = _mix64(np.uint64(key) + np.arange((L + 1) // 2, dtype=np.uint64))
    u = np.empty(2 * len(x), np.uint64)
    u[0::2] = x & np.uint64(0xFFFFFFFF)
    u[1::2] = x >> np.uint64(32)
    u = u[:L]
    t = [min(4294967295, int(np.floor(f * 4294967296.0))) for f in (gc / 2, gc, gc + (1 - gc) / 2)]
    return np.where(u < t[0], 1, np.where(u < t[1], 3, np.where(u < t[2], 0, 4))).astype(np.int8)


def test_sw_random_pairs_match_oracle(gpu, oracle_lib):
    """Random fragment pairs generated on the device: scores equal the
    oracle's SmithWaterman on the same fragments regenerated on the host."""
    with np.errstate(over="ignore"):
        pts = [(0.3, 38, 2000), (0.5, 700, 41), (0.62, 333, 333), (0.45, 1999, 1500), (0.7, 64, 65)]
        sc = (2, -3, -4, -2)
        per = 64
        seeds = np.array([12345 + 1000 * k for k in range(len(pts))], np.uint64)
        print("sw random", pts, flush=True)
        got = gpu.sw_random_max(sc, [p[0] for p in pts], [p[1] for p in pts], [p[2] for p in pts], seeds, per)
        seqs, o1, l1, o2, l2 = [], [], [], [], []
        off = 0
        for k, (g, L1, L2) in enumerate(pts):
            for r in range(per):
                a = _random_fragment(_mix64(seeds[k] + np.uint64(2 * r)), L1, g)
                b = _random_fragment(_mix64(seeds[k] + np.uint64(2 * r + 1)), L2, g)
                seqs += [a, b]
                o1.append(off); l1.append(L1); off += L1
                o2.append(off); l2.append(L2); off += L2
        want = bind_sw_max(np.concatenate(seqs), o1, l1, o2, l2, sc)
        assert np.array_equal(got.ravel(), want)
        # the generator's composition
        a = np.concatenate([_random_fragment(_mix64(np.uint64(7 + r)), 5000, 0.62) for r in range(8)])
        assert abs(np.isin(a, (1, 3)).mean() - 0.62) < 0.01


def bind_sw_max(seqs, o1, l1, o2, l2, sc):
    from oracle import bind
    return bind.sw_max_batch(seqs, np.array(o1, np.int64), np.array(l1, np.int32), np.array(o2, np.int64),
                             np.array(l2, np.int32), sc, nthreads=8)


@pytest.mark.parametrize("sc", synth.scoring_settings())
def test_sw_random_config3_scorings_match_oracle(gpu, sc):
    """The six scoring settings of the permutation benchmark at its shapes:
    one pair per point (every lane its own lengths and composition), lengths
    log-uniform on 38-2000 and GC uniform on 0.30-0.70, plus a few points of
    64 pairs; scores equal the oracle's SmithWaterman on the same fragments
    regenerated on the host."""
    rng = np.random.default_rng(1000 + sum(abs(x) * 7 ** k for k, x in enumerate(sc)))
    npts = 384
    l1 = synth.log_uniform_int(rng, 38, 2000, npts).astype(np.int32)
    l2 = synth.log_uniform_int(rng, 38, 2000, npts).astype(np.int32)
    gc = rng.uniform(0.30, 0.70, npts)
    seeds = rng.integers(0, 2**62, npts).astype(np.uint64)
    print("sw random config-3 shapes", sc, npts, flush=True)
    got = gpu.sw_random_max(sc, gc, l1, l2, seeds, 1).ravel()
    got64 = gpu.sw_random_max(sc, gc[:4], l1[:4], l2[:4], seeds[:4], 64).ravel()
    with np.errstate(over="ignore"):
        seqs, o1, ln1, o2, ln2 = [], [], [], [], []
        off = 0

        def add(a, b):
            nonlocal off
            seqs.extend([a, b])
            o1.append(off); ln1.append(len(a)); off += len(a)
            o2.append(off); ln2.append(len(b)); off += len(b)

        for k in range(npts):
            add(_random_fragment(_mix64(seeds[k]), int(l1[k]), gc[k]),
                _random_fragment(_mix64(seeds[k] + np.uint64(1)), int(l2[k]), gc[k]))
        for k in range(4):
            for r in range(64):
                add(_random_fragment(_mix64(seeds[k] + np.uint64(2 * r)), int(l1[k]), gc[k]),
                    _random_fragment(_mix64(seeds[k] + np.uint64(2 * r + 1)), int(l2[k]), gc[k]))
    want = bind_sw_max(np.concatenate(seqs), o1, ln1, o2, ln2, sc)
    assert np.array_equal(got, want[:npts])
    assert np.array_equal(got64, want[npts:])
    assert got.max() > 0


def test_sw_bas_matches_reference_body():
    """dcns_sw_bas_batch against the reference's SmithWaterman_bas body
    compiled as its own translation unit (tests/golden/sw_bas.json): maxScore
    and end positions for every case (N against N scores -1, gap open + k *
    extend), and with reverseAlignment the reference's traceback walk: the
    cigar where the walk ends, the zero-cell stop ("line 110"), and the cases
    where it never ends."""
    import json
    import os
    import dcns_amd
    gold = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sw_bas.json")))
    ctx = dcns_amd.Context(0)
    try:
        for rev in (0, 1):
            for key in sorted({(c["open"], c["ext"]) for c in gold if c["reverse"] == rev}):
                cs = [c for c in gold if c["reverse"] == rev and (c["open"], c["ext"]) == key]
                pairs = [(synth.encode(c["seq1"]), synth.encode(c["seq2"])) for c in cs]
                print("sw_bas", rev, key, len(cs), flush=True)
                got = ctx.sw_bas(pairs, open=key[0], ext=key[1], reverse=bool(rev))
                for c, g in zip(cs, got):
                    res, cig, st = (g, None, 0) if not rev else g
                    if c["never_ends"]:
                        assert st == 2, (c, g)
                        if c["max"] is not None:
                            assert res[0] == c["max"], (c, g)
                        continue
                    assert res == (c["max"], c["end1"], c["end2"]), (c, g)
                    if rev:
                        assert cig == c["cigar"], (c, g)
                        assert st == (1 if c["line110"] else 0), (c, g)
    finally:
        ctx.close()
    assert sum(c["never_ends"] for c in gold) > 20 and sum(1 for c in gold if c["reverse"] and not c["never_ends"]) > 10


def test_sw_random_packed_overflow_rerun_and_lanes(gpu):
    """The K1 launch shapes (tuning k1_lanes): one int32 pair per lane, two
    packed f16 pairs per lane, one pair per wave, and the routed mix give the
    same maxima, on shapes where some maxima pass 2048 (GC 1.0 with a
    linear-regime scoring, so packed pairs take the int32 re-run), plus a
    scoring the packed kernel does not take (f16 match 9 has a nonzero low
    byte); the overflowing ones also against the oracle."""
    pts = [(1.0, 2000, 2000), (1.0, 1900, 1700), (0.98, 2000, 1999), (0.5, 300, 2000), (0.35, 38, 39)]
    seeds = np.array([777 + 31 * k for k in range(len(pts))], np.uint64)
    gc = [p[0] for p in pts]
    l1 = [p[1] for p in pts]
    l2 = [p[2] for p in pts]
    old = gpu.get_tuning()
    try:
        for sc in ((3, -1, -2, -1), (9, -3, -4, -2), (2, -3, -4, -2)):
            got = {}
            for mode in (1, 2, 3, 0):  # int32 lanes, packed lanes, waves, the routed mix
                print("sw random lanes", sc, mode, flush=True)
                gpu.set_tuning(k1_lanes=mode)
                got[mode] = gpu.sw_random_max(sc, gc, l1, l2, seeds, 70)
            for mode in (2, 3, 0):
                assert np.array_equal(got[mode], got[1]), (sc, mode)
            if sc == (3, -1, -2, -1):
                assert got[2].max() >= 2048 and got[2].min() >= 0
            # the same shapes again with other seeds: the cached launch plan
            # (after a call whose overflowing pairs re-ran) against int32 lanes
            seeds2 = seeds + np.uint64(99991)
            gpu.set_tuning(k1_lanes=0)
            a = gpu.sw_random_max(sc, gc, l1, l2, seeds2, 70)
            b = gpu.sw_random_max(sc, gc, l1, l2, seeds2, 70)
            gpu.set_tuning(k1_lanes=1)
            assert np.array_equal(a, gpu.sw_random_max(sc, gc, l1, l2, seeds2, 70)), sc
            assert np.array_equal(a, b), sc
    finally:
        gpu.set_tuning(**old)
    # the linear-regime pairs (maxima past 2048) against the oracle, packed lanes
    gpu.set_tuning(k1_lanes=2)
    try:
        got = gpu.sw_random_max((3, -1, -2, -1), gc[:1], l1[:1], l2[:1], seeds[:1], 6).ravel()
    finally:
        gpu.set_tuning(**old)
    with np.errstate(over="ignore"):
        seqs, o1, ln1, o2, ln2 = [], [], [], [], []
        off = 0
        for r in range(6):
            a = _random_fragment(_mix64(seeds[0] + np.uint64(2 * r)), l1[0], gc[0])
            b = _random_fragment(_mix64(seeds[0] + np.uint64(2 * r + 1)), l2[0], gc[0])
            seqs += [a, b]
            o1.append(off); ln1.append(len(a)); off += len(a)
            o2.append(off); ln2.append(len(b)); off += len(b)
    want = bind_sw_max(np.concatenate(seqs), o1, ln1, o2, ln2, (3, -1, -2, -1))
    assert np.array_equal(got, want)
    assert got.max() >= 2048

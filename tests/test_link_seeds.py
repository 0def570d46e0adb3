"""link_seeds on the host (dcns_link_seeds, the block-evaluated pass) against
the oracle's one-seed-at-a-time restatement and the compiled reference's own
link_seeds (impl/findSimilarFragmentsForPairedSequence4.cpp:288-344), on seed
lists that take several merge passes: same-length seeds along diagonals
(merges), seeds inside others (absorbed), seeds sharing a start or an end
pair (equivalent under Seed's comparator, so the set keeps one), and the
oracle's unlinked seed sets of planted pairs.  No device is needed."""
import numpy as np
import pytest

import dcns_amd
from dcns_amd import synth


def _diagonal_seeds(rng, n, span=4000, diagonals=12):
    out = []
    diags = rng.integers(-span // 4, span // 4, diagonals)
    while len(out) < n:
        d = int(rng.choice(diags))
        kind = rng.random()
        s1 = int(rng.integers(0, span))
        ln = int(rng.choice([20, 24, 30, 38, 40])) if kind < 0.8 else int(rng.integers(5, 120))
        if kind < 0.5:  # a run of same-length seeds one step apart: merge chains
            step = int(rng.integers(1, 12))
            for k in range(int(rng.integers(2, 8))):
                a = s1 + k * step
                out.append((a, a + ln, a + d, a + d + ln))
        elif kind < 0.7:  # a seed and seeds inside it
            out.append((s1, s1 + ln + 30, s1 + d, s1 + d + ln + 30))
            for _ in range(int(rng.integers(1, 4))):
                o = int(rng.integers(0, 25))
                out.append((s1 + o, s1 + o + ln // 2, s1 + d + o, s1 + d + o + ln // 2))
        elif kind < 0.8:  # off-diagonal neighbours (gapped seeds)
            g = int(rng.integers(1, 6))
            out.append((s1, s1 + ln, s1 + d, s1 + d + ln + g))
        else:
            out.append((s1, s1 + ln, s1 + d, s1 + d + ln))
    out = [s for s in out if min(s) >= 0]
    order = rng.permutation(len(out))
    return [out[k] for k in order]


@pytest.mark.parametrize("seed,n", [(1, 50), (2, 400), (3, 3000), (4, 20000)])
def test_link_seeds_vs_oracle_and_reference(oracle_lib, seed, n):
    from oracle import bind
    rng = np.random.default_rng(seed)
    seeds = _diagonal_seeds(rng, n)
    got = dcns_amd.link_seeds(seeds)
    want = oracle_lib.link_seeds(seeds)
    assert got == want
    assert dcns_amd.link_seeds(seeds, insertion_order=True) == oracle_lib.link_seeds(seeds, insertion_order=True)
    assert len(got) < len(set(seeds))  # merges and absorptions happened
    ref = bind.reference()
    if ref is not None:
        assert ref.link_seeds(seeds) == want


def test_link_seeds_edge_lists(oracle_lib):
    cases = [[], [(5, 45, 7, 47)], [(5, 45, 7, 47)] * 3,
             [(0, 40, 0, 40), (0, 30, 0, 50), (10, 40, 12, 40)],  # shared start pair, shared end pair
             [(0, 40, 0, 40), (1, 41, 1, 41), (2, 42, 2, 42), (3, 43, 3, 43), (4, 44, 4, 44)],
             [(0, 40, 100, 140), (50, 90, 0, 40), (60, 100, 160, 200)]]
    for seeds in cases:
        assert dcns_amd.link_seeds(seeds) == oracle_lib.link_seeds(seeds), seeds


def test_link_seeds_of_unlinked_sets(oracle_lib):
    """The unlinked seed sets the oracle's window loop builds, linked by the
    product, equal the oracle's linked sets."""
    for k in range(3):
        a, b = synth.one_pair(seed=7100 + k, L1=6000, L2=5000, nfrag=12)[:2]
        unlinked = oracle_lib.seed_set(a, b, linked=False)
        assert dcns_amd.link_seeds(unlinked) == oracle_lib.seed_set(a, b, linked=True), k


@pytest.mark.parametrize("seed,n,span", [(11, 100000, 100000), (12, 100000, 30000), (14, 80000, 4000)])
def test_link_seeds_large_lists(oracle_lib, seed, n, span):
    """Lists of 65,536 seeds and more link on all host threads, in chunks cut
    where no earlier seed reaches the next one, with consumed seeds across a
    chunk's end replayed in order: the result equals the one-seed-at-a-time
    restatement's, sparse (many chunks), medium and dense (few, long)."""
    from oracle import bind
    rng = np.random.default_rng(seed)
    seeds = _diagonal_seeds(rng, n, span=span, diagonals=40)
    assert len(set(seeds)) >= 65536
    want = oracle_lib.link_seeds(seeds)
    assert dcns_amd.link_seeds(seeds) == want
    assert dcns_amd.link_seeds(seeds, insertion_order=True) == oracle_lib.link_seeds(seeds, insertion_order=True)
    ref = bind.reference()
    if ref is not None:
        assert ref.link_seeds(seeds) == want

"""The CPU oracle against the compiled reference, live, on fresh random
inputs (skipped where oracle/_ref has not been built)."""
import numpy as np
import pytest

from dcns_amd import synth


def test_k1_live(oracle_lib, ref_lib):
    rng = np.random.default_rng(41)
    for t in range(400):
        sc = (int(rng.integers(1, 4)), -int(rng.integers(1, 5)), -int(rng.integers(1, 7)), -int(rng.integers(1, 4)))
        a = rng.choice([0, 1, 2, 3, 4], int(rng.integers(0, 150)), p=[.24, .24, .04, .24, .24]).astype(np.int8)
        b = rng.choice([0, 1, 2, 3, 4], int(rng.integers(0, 70)), p=[.24, .24, .04, .24, .24]).astype(np.int8)
        x, y = oracle_lib.sw(a, b, sc), ref_lib.sw(a, b, sc)
        assert x[0] == y[0] and (x[0] == 0 or x == y)
        assert oracle_lib.sw(a, b, sc, True) == ref_lib.sw(a, b, sc, True)


def test_extension_live(oracle_lib, ref_lib):
    rng = np.random.default_rng(42)
    for t in range(300):
        sc = (2, -3, -4, -2) if t % 2 else (int(rng.integers(1, 4)), -int(rng.integers(1, 5)),
                                             -int(rng.integers(1, 7)), -int(rng.integers(1, 4)))
        a = rng.choice([0, 1, 3, 4], int(rng.integers(1, 400))).astype(np.int8)
        b = synth.mutate(rng, a, .15, .04)[:int(rng.integers(1, 400))]
        if len(b) == 0:
            continue
        w, xd = int(rng.integers(1, 15)), int(rng.integers(5, 40))
        assert oracle_lib.xextend(a, b, sc, xd, w) == ref_lib.xextend(a, b, sc, xd, w)
        assert oracle_lib.xextend_tb(a, b, sc, xd, w) == ref_lib.xextend_tb(a, b, sc, xd, w)
        sc2 = sc + (-int(rng.integers(10, 50)), -int(rng.integers(0, 2)))
        z, w2 = int(rng.integers(10, 80)), int(rng.integers(3, 400))
        assert oracle_lib.semiglobal(a, b, sc2, z, w2) == ref_lib.semiglobal(a, b, sc2, z, w2)


def test_pipeline_live(oracle_lib, ref_lib):
    rng = np.random.default_rng(43)
    for t in range(8):
        a, b = synth.one_pair(seed=900 + t, L1=int(rng.integers(200, 2500)), L2=int(rng.integers(60, 2500)),
                              nfrag=int(rng.integers(0, 10)))
        assert oracle_lib.seed_set(a, b) == ref_lib.seed_set(a, b)
        assert oracle_lib.seed_set(a, b, linked=True) == ref_lib.seed_set(a, b, linked=True)
        assert oracle_lib.find_similar(a, b) == ref_lib.find_similar(a, b)


@pytest.mark.parametrize("kw", [
    dict(sc=(2, -3, -4, -2, -45, 0), zdrop=50, bandwidth=20000),
    dict(sc=(2, -3, -4, -2, -45, 0), zdrop=50, bandwidth=120),
    dict(sc=(2, -3, -4, -2, -20, -1), zdrop=30, bandwidth=20000),
    dict(sc=(2, -3, -4, -2, -45, 0), zdrop=10, bandwidth=3),
])
def test_two_piece_live(oracle_lib, ref_lib, kw):
    rng = np.random.default_rng(44 + kw["bandwidth"])
    for t in range(3):
        a, b = synth.one_pair(seed=950 + t, L1=int(rng.integers(300, 2500)), L2=int(rng.integers(200, 2000)),
                              nfrag=int(rng.integers(1, 8)))
        assert oracle_lib.find_similar_2gaps(a, b, **kw) == ref_lib.find_similar_2gaps(a, b, **kw)

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

"""K1 semantics pinned without the reference: exhaustive path enumeration on
tiny inputs.  The reference's striped SmithWaterman forbids a vertical gap
(along the reference) right after a horizontal one (along the read), because
its E update reads H before the horizontal-gap fix-up
(sequenceAlignment3.cpp:283-294 vs :309-322); the scalar textbook path does
not.  Enumerating every local alignment path decides both the maximum and
its first (ref-major) end cell independently of any DP formulation."""
import itertools

import numpy as np
import pytest


def brute(ref, read, sc, variant):
    match, mism, open_, ext = sc
    L1, L2 = len(ref), len(read)

    def s(a, b):
        if a == 2 or b == 2:
            return 0
        return match if a == b else mism

    best = {}  # grid point (i, j) -> best score of a non-empty path ending there

    def walk(i, j, score, last):
        if last is not None:
            if best.get((i, j), -10**9) < score:
                best[(i, j)] = score
        if i < L1 and j < L2:
            walk(i + 1, j + 1, score + s(ref[i], read[j]), "D")
        # a run may continue (extend) or restart (open) after a run of its own kind
        if i < L1 and not (variant and last == "H"):
            walk(i + 1, j, score + (max(ext, open_) if last == "V" else open_), "V")
        if j < L2:
            walk(i, j + 1, score + (max(ext, open_) if last == "H" else open_), "H")

    for i0 in range(L1 + 1):
        for j0 in range(L2 + 1):
            walk(i0, j0, 0, None)
    mx, pos = 0, (-1, -1)
    for i in range(1, L1 + 1):
        for j in range(1, L2 + 1):
            v = max(0, best.get((i, j), 0))
            if v > mx:
                mx, pos = v, (i - 1, j - 1)
    return (mx, *pos)


SCORINGS = [(2, -3, -4, -2), (1, -4, -2, -1), (3, -1, -5, -1), (2, -2, -1, -1), (2, -3, -1, -3)]


@pytest.mark.parametrize("sc", SCORINGS)
def test_k1_variant_by_enumeration(oracle_lib, sc):
    rng = np.random.default_rng(sum(sc) + 100)
    for _ in range(60):
        a = rng.choice([0, 1, 2, 3, 4], int(rng.integers(1, 5)), p=[.23, .23, .08, .23, .23]).astype(np.int8)
        b = rng.choice([0, 1, 2, 3, 4], int(rng.integers(1, 5)), p=[.23, .23, .08, .23, .23]).astype(np.int8)
        assert oracle_lib.sw(a, b, sc) == brute(a, b, sc, True), (a, b)
        assert oracle_lib.sw(a, b, sc, textbook=True) == brute(a, b, sc, False), (a, b)


def test_variant_and_textbook_agree_on_every_cell(oracle_lib):
    """Forbidding a vertical gap right after a horizontal one never changes a
    cell's best score: the two runs can be swapped (vertical first) without
    changing the score or the end cell.  So the striped and textbook paths of
    SmithWaterman report the same maximum and position, under any scoring;
    exhaustive check on all length-3 inputs under cheap gaps."""
    sc = (3, -5, -1, -1)
    for a in itertools.product([0, 1, 3], repeat=3):
        for b in itertools.product([0, 1, 3], repeat=3):
            x, y = np.array(a, np.int8), np.array(b, np.int8)
            v, t = oracle_lib.sw(x, y, sc), oracle_lib.sw(x, y, sc, textbook=True)
            assert v == t == brute(x, y, sc, True) == brute(x, y, sc, False)


def test_default_scoring_variant_equals_textbook(oracle_lib):
    """Under cut1Gap defaults an adjacent gap pair never beats mismatches, so
    the two recurrences agree on scores."""
    rng = np.random.default_rng(1)
    for _ in range(300):
        a = rng.choice([0, 1, 3, 4], int(rng.integers(1, 60))).astype(np.int8)
        b = rng.choice([0, 1, 3, 4], int(rng.integers(1, 40))).astype(np.int8)
        assert oracle_lib.sw(a, b)[0] == oracle_lib.sw(a, b, textbook=True)[0]


def test_identical_and_all_n(oracle_lib):
    rng = np.random.default_rng(2)
    a = rng.choice([0, 1, 3, 4], 300).astype(np.int8)
    assert oracle_lib.sw(a, a) == (600, 299, 299)
    n = np.full(80, 2, np.int8)
    assert oracle_lib.sw(n, a[:40])[0] == 0
    assert oracle_lib.seed_set(n, n) == []

"""The bound behind the rescans' level-0 tails, checked by brute force on the
window scan's recurrence (the form every scan kernel evaluates): a DP
started from zeros at row a equals, cell by cell, the DP started at row 0
on every row from a + D on, D = wm + ceil(window * match / g) with
g = -max(open, ext) and wm the columns the kernels evaluate (the window
padded with N columns, which score 0).  A path from above row a to such a
row enters at most wm rows diagonally, gains at most match on each real
column and pays at least g for each row it enters by a vertical gap step,
so its score cannot be positive.  The cases plant copies of the window before the
rescan's start (high scores that would carry over) and run scorings whose
gap opening is cheaper or dearer than an extension."""
import numpy as np
import pytest

from test_rescan_trim import sw_rows


def tail_rows(window, wm, match, open_, ext):
    g = -max(open_, ext)
    return wm + (window * match + g - 1) // g


def _case(rng, L1, W, a, real=None):
    ref = rng.choice(np.array([0, 1, 3, 4], np.int8), L1)
    win = rng.choice(np.array([0, 1, 3, 4], np.int8), W)
    if real is not None:
        win[real:] = 2  # a partial window: N columns past its end
    # a copy of the window ending just above row a, split by a gap run, so
    # that high scores cross the start row along several paths
    k = int(rng.integers(W // 3, W - 3))
    gap = int(rng.integers(0, 12))
    top = a - k
    if top >= 0:
        ref[top:a] = win[:k]
        ref[a + gap:a + gap + W - k] = win[k:]
    if rng.random() < 0.2:
        ref[rng.random(L1) < 0.05] = 2  # N bases
    return ref, win


@pytest.mark.parametrize("sc,real", [((2, -3, -4, -2), None), ((1, -1, -2, -1), None), ((2, -3, -1, -3), None),
                                     ((3, -4, -6, -3), None), ((2, -3, -4, -2), 20), ((2, -3, -1, -3), 9)])
def test_rescan_equals_level0_from_the_tail_row(sc, real):
    match, mismatch, open_, ext = sc
    rng = np.random.default_rng((abs(hash(sc)) + (real or 0)) % 2**32)
    W = 38
    D = tail_rows(real or W, W, match, open_, ext)
    differing_before = 0
    for case in range(12):
        L1 = int(rng.integers(D + 60, D + 160))
        a = int(rng.integers(W, L1 - D - 20))
        ref, win = _case(rng, L1, W, a, real)
        full = sw_rows(ref, win, 0, L1, match, mismatch, open_, ext)
        sub = sw_rows(ref, win, a, L1, match, mismatch, open_, ext)
        assert (sub[D:] == full[a + D:]).all(), (sc, case, a)
        differing_before += int((sub[:D] != full[a:a + D]).any())
    # the planted copies make the rows before a + D differ: the equality is
    # not vacuous
    assert differing_before >= 6


def test_planted_mistake_a_short_tail_differs():
    """A tail of 30 rows (fewer than the window) is not enough: a copy of the
    window split by a gap run across the start row carries its score past
    it, and the check above would fail."""
    rng = np.random.default_rng(99)
    W, D = 38, 30
    found = False
    for case in range(20):
        L1, a = 260, 120
        ref, win = _case(rng, L1, W, a)
        full = sw_rows(ref, win, 0, L1)
        sub = sw_rows(ref, win, a, L1)
        if (sub[D:] != full[a + D:]).any():
            found = True
            break
    assert found

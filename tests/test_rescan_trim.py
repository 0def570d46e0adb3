"""The property the seed recursion's rescan trim rests on, checked by brute
force on the textbook SmithWaterman recurrence of the window scan (the form
every scan kernel evaluates; the brute-force path enumeration of
test_oracle_bruteforce.py ties it to the reference's striped form): a DP
started from zeros at row a never exceeds, cell by cell, the DP started at
row 0, so every trajectory entry of a rescan over rows [a, b) lies on a row
where the full scan's row maximum reached min_score.  The cases plant
repeated copies of the window so that rescans do find seeds."""
import numpy as np


def sw_rows(ref, win, a, b, match=2, mismatch=-3, open_=-4, ext=-2):
    """H of rows [a, b) of the window scan started from zeros at row a."""
    W = len(win)
    H = np.zeros(W, np.int64)
    E = np.zeros(W, np.int64)
    out = np.zeros((b - a, W), np.int64)
    for i in range(a, b):
        diag, f = 0, 0
        for j in range(W):
            c, q = ref[i], win[j]
            s = 0 if (c == 2 or q == 2) else (match if c == q else mismatch)
            h = max(diag + s, E[j], f, 0)
            ho = h + open_
            E[j] = max(E[j] + ext, ho, 0)
            f = max(ho, f + ext, 0)
            diag = H[j]
            H[j] = h
            out[i - a, j] = h
    return out


def trajectory_rows(Hm, a, min_score):
    best, rows = min_score - 1, []
    for r in range(Hm.shape[0]):
        m = int(Hm[r].max())
        if m > best:
            best = m
            rows.append(a + r)
    return rows


def test_rescan_never_exceeds_level0_and_entries_lie_on_strong_rows():
    rng = np.random.default_rng(2024)
    rescans_with_entries = 0
    for case in range(40):
        L1, W = int(rng.integers(120, 260)), 38
        ref = rng.choice(np.array([0, 1, 3, 4], np.int8), L1)
        win = rng.choice(np.array([0, 1, 3, 4], np.int8), W)
        # plant 1-3 mutated copies of the window (repeats: rescans find seeds)
        for _ in range(int(rng.integers(1, 4))):
            p = int(rng.integers(0, L1 - W))
            cp = win.copy()
            mut = rng.random(W) < 0.08
            cp[mut] = rng.choice(np.array([0, 1, 3, 4], np.int8), int(mut.sum()))
            ref[p:p + W] = cp
        if case % 5 == 0:
            ref[rng.random(L1) < 0.05] = 2  # N bases
        full = sw_rows(ref, win, 0, L1)
        strong = {i for i in range(L1) if full[i].max() >= 40}
        for a in sorted(set(int(x) for x in rng.integers(1, L1 - 1, 6))):
            sub = sw_rows(ref, win, a, L1)
            assert (sub <= full[a:]).all(), (case, a)
            rows = trajectory_rows(sub, a, 40)
            assert set(rows) <= strong, (case, a, rows)
            rescans_with_entries += bool(rows)
    assert rescans_with_entries >= 10


def test_planted_mistake_breaks_the_property():
    """A recurrence without the 0 floor is not monotone in the start row:
    the check above must then fail on some case (the test detects a wrong
    recurrence rather than passing vacuously)."""
    rng = np.random.default_rng(7)

    def sw_nofloor(ref, win, a, b):
        W = len(win)
        H = np.zeros(W, np.int64)
        E = np.full(W, -10**6, np.int64)
        out = np.zeros((b - a, W), np.int64)
        for i in range(a, b):
            diag, f = (0 if i == a else -10**6), -10**6
            for j in range(W):
                s = 2 if ref[i] == win[j] else -3
                h = max(diag + s, E[j], f)
                E[j] = max(E[j] - 2, h - 4)
                f = max(h - 4, f - 2)
                diag = H[j]
                H[j] = h
                out[i - a, j] = h
        return out

    broke = False
    for _ in range(10):
        ref = rng.choice(np.array([0, 1, 3, 4], np.int8), 80)
        win = rng.choice(np.array([0, 1, 3, 4], np.int8), 20)
        full = sw_nofloor(ref, win, 0, 80)
        sub = sw_nofloor(ref, win, 30, 80)
        if not (sub <= full[30:]).all():
            broke = True
            break
    assert broke

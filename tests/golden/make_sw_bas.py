"""Writes sw_bas.json: the reference's SmithWaterman_bas body
(impl/SmithWaterman_bas.cpp:59-201, which the reference's own build leaves
out) compiled as its own translation unit (oracle/_ref/swbas_ref, built by
`make -C oracle ref`) on small pairs with N, gap settings and both
reverseAlignment values: maxScore, the end positions and the cigar, the
"line 110" message when the traceback stops at a zero cell, or "never ends"
when the traceback loop does not return within 2 s (its walk stops moving and
the process spins).  Only data is written:

    make -C oracle ref && python tests/golden/make_sw_bas.py
"""
import json
import multiprocessing as mp
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "swbas_ref")
sys.path.insert(0, ROOT)

from dcns_amd import synth  # noqa: E402


def cases():
    rng = np.random.default_rng(20261204)
    out = []
    gaps = [(-2, -1), (-4, -2), (-1, -1), (0, -1), (-3, 0), (-5, -3)]
    for t in range(240):
        L1, L2 = int(rng.integers(0, 48)), int(rng.integers(0, 40))
        p_n = [0.0, 0.05, 0.3][t % 3]
        pn = [(1 - p_n) / 4] * 2 + [p_n] + [(1 - p_n) / 4] * 2
        a = rng.choice([0, 1, 2, 3, 4], L1, p=pn).astype(np.int8)
        b = rng.choice([0, 1, 2, 3, 4], L2, p=pn).astype(np.int8)
        if L1 > 12 and L2 > 8 and t % 2:
            k = int(rng.integers(0, L1 - 8))
            m = synth.mutate(rng, a[k:k + L2], 0.15, 0.05)[:L2]
            b[:len(m)] = m
        o, e = gaps[t % len(gaps)]
        out.append((a, b, o, e, (t // 2) % 2))
    # single matches and other short walks the traceback can finish
    for x in ([0], [1, 2], [3, 2, 2], [2, 2, 4]):
        for y in ([0], [2, 0], [2, 2, 3], [4, 2, 2, 1]):
            out.append((np.array(x, np.int8), np.array(y, np.int8), -2, -1, 1))
    # longer pairs, forward only
    for t in range(12):
        a = synth.random_seq(rng, int(rng.integers(150, 400)))
        b = synth.mutate(rng, a[20:300], 0.1, 0.03)
        b[::17] = 2
        out.append((a, b, -4, -2, 0))
    return out


def run_ref(case):
    a, b, o, e, rev = case
    text = f"{len(a)} {len(b)} {o} {e} {rev}\n{' '.join(map(str, a))}\n{' '.join(map(str, b))}\n"
    r = subprocess.run(["bash", "-o", "pipefail", "-c", f"timeout -s KILL 2 {REF} 2>/tmp/_bas_err_$$ | tail -c 4096;"
                        " rc=$?; cat /tmp/_bas_err_$$ >&2; rm -f /tmp/_bas_err_$$; exit $rc"],
                       input=text, capture_output=True, text=True)
    if r.returncode == 137:  # killed at the time limit: the walk never ended
        first = r.stdout.splitlines()[0] if r.stdout else ""  # the body prints maxScore before its walk
        return {"never_ends": True, "max": int(first.split()[1]) if first.startswith("maxScore ") else None}
    if r.returncode != 0:
        raise RuntimeError(f"the reference failed ({r.returncode}) on {text[:80]!r}")
    last = [ln for ln in r.stdout.splitlines() if ln.startswith("RESULT")][-1].split()
    mx, e1, e2, n = (int(v) for v in last[1:5])
    return {"never_ends": False, "max": mx, "end1": e1, "end2": e2, "cigar": [int(v) for v in last[5:5 + n]],
            "line110": "line 110" in r.stderr}


def main():
    cs = cases()
    with mp.Pool(8) as pool:
        res = pool.map(run_ref, cs, chunksize=1)
    out = [{"seq1": synth.to_text(a), "seq2": synth.to_text(b), "open": o, "ext": e, "reverse": rev, **r}
           for (a, b, o, e, rev), r in zip(cs, res)]
    with open(os.path.join(HERE, "sw_bas.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
    ends = sum(1 for x in out if x["reverse"] and not x["never_ends"])
    print(f"sw_bas.json: {len(out)} cases, {sum(x['never_ends'] for x in out)} never end, "
          f"{ends} reverse cases end ({sum(1 for x in out if x.get('line110'))} at line 110)")


if __name__ == "__main__":
    main()

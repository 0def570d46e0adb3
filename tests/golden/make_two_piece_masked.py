"""Writes two_piece_masked.json: findSimilarFragmentsForPairedSequence
overload 4 (cut2Gaps defaults: second gap piece -45 / 0, z-drop 50, band
20000) on a long-masked pair of the benchmark's shape at a size the oracle
finishes in about a minute and a half: 20 kb regions at GC 0.45 with 30%
lowercase-n runs (stripped, as cut2Gaps strips them), 8 copies of a 3 kb
repeat at 10% divergence and 20 planted fragments 25-2,000 bp apart
(dcns_amd.synth.long_masked(seed=88, L=20000, te_copies=8, nfrag=20)).  The
fragments and reference cell counts come from the oracle; the fragments are
checked against the compiled reference (oracle/_ref).  Only data is written:

    make -C oracle oracle ref && python tests/golden/make_two_piece_masked.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from dcns_amd import synth  # noqa: E402
from oracle import bind  # noqa: E402

ARGS = dict(seed=88, L=20000, te_copies=8, te_len=3000, nfrag=20)


def pair():
    t1, t2 = synth.long_masked(**ARGS)
    return synth.encode(t1.replace("n", "")), synth.encode(t2.replace("n", ""))


def main():
    a, b = pair()
    st = {}
    want = bind.oracle().find_similar_2gaps(a, b, stats=st)
    ref = bind.reference()
    if ref is None or ref.find_similar_2gaps(a, b) != want:
        sys.exit("the compiled reference is missing or differs from the oracle")
    out = {"workload": "dcns_amd.synth.long_masked(" + ", ".join(f"{k}={v}" for k, v in ARGS.items()) +
                       "), n stripped, cut2Gaps defaults",
           "len1": len(a), "len2": len(b), "stats": st,
           "fragments": [list(f[:7]) + [list(f[7])] for f in want]}
    with open(os.path.join(HERE, "two_piece_masked.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print(f"two_piece_masked.json: {len(want)} fragments, {st}")


if __name__ == "__main__":
    main()

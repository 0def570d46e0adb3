"""Autotuner of the seed phase's launch shapes (dcns_tuning) on a staged batch.

The reference runs one pair per process, so it has nothing to tune; here
thousands of pairs share one launch per recursion level, and how they share
it is a choice whose best value depends on the batch:

  * level0_mode / level0_waves: the level-0 window scan as one workgroup per
    128-window task, or as persistent waves popping the length-sorted task
    list (and how many per CU);
  * wave_tasks / wave_rows: how many of the longest later-level rescans run
    one per wave (64 lanes on one task) instead of one per lane;
  * graph_levels: how many recursion levels are replayed as one captured HIP
    graph without host reads between them;
  * xdrop_lane_rows: how many rows an x-drop pass may run on the
    one-seed-per-lane kernels before its seed moves to one wave.

Results never depend on these settings, only the time does.  `tune` measures
the whole findSimilarFragmentsForPairedSequence step on the staged batch,
one setting at a time (coordinate descent from the defaults, the best of
`reps` runs per candidate), and returns the best setting with every
measurement.  The table of best settings per batch shape is kept in
data/seed_tuning.json:

    python -m dcns_amd.autotune [--pairs 4096] [--write]
"""
import json
import math
import os
import time

import numpy as np

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "seed_tuning.json")

DEFAULT = {"level0_mode": 1, "level0_waves": 8, "wave_tasks": 1024, "wave_rows": 4096, "graph_levels": 4,
           "pool_min": 0, "xdrop_lane_rows": 512}
SPACE = [("level0_mode", [1, 2]), ("level0_waves", [4, 8]), ("wave_tasks", [256, 1024, 2048]),
         ("wave_rows", [2048, 4096, 8192]), ("graph_levels", [-1, 4, 8]),
         ("xdrop_lane_rows", [256, 512, 1024, 4096, -1])]


def shape_key(len1, len2):
    """Batch shape: pair count (power of two) and mean lengths (powers of
    two in kb), the quantities the launch shapes depend on."""
    n = len(len1)
    m1 = float(np.mean(len1)) / 1000 if n else 0.0
    m2 = float(np.mean(len2)) / 1000 if n else 0.0
    b = lambda x: int(round(math.log2(max(x, 1e-3))))  # noqa: E731
    return f"pairs2^{b(n)}_ref2^{b(m1)}kb_qry2^{b(m2)}kb"


def _time_step(ctx, params, reps):
    best = float("inf")
    for _ in range(reps):
        t = time.perf_counter()
        ctx.find_similar_staged(params, collect=False)
        best = min(best, time.perf_counter() - t)
    return best * 1e3


def tune(ctx, params=None, reps=2, log=None):
    """Coordinate descent over SPACE on the context's staged batch.  Returns
    (best setting, rows of every measurement), leaves the best one set."""
    cur = dict(DEFAULT)
    rows = []
    ctx.set_tuning(**cur)
    ctx.find_similar_staged(params, collect=False)  # pools sized, graph captured
    best_ms = _time_step(ctx, params, reps)
    rows.append({**cur, "ms_per_step": round(best_ms, 2)})
    for field, values in SPACE:
        if field == "level0_waves" and cur["level0_mode"] != 2:
            continue
        for v in values:
            if v == cur[field]:
                continue
            cand = {**cur, field: v}
            ctx.set_tuning(**cand)
            ctx.find_similar_staged(params, collect=False)  # capture / first-use costs out of the timing
            ms = _time_step(ctx, params, reps)
            rows.append({**cand, "ms_per_step": round(ms, 2)})
            if log:
                log(f"autotune: {field}={v}: {ms:.1f} ms/step (best so far {best_ms:.1f})")
            if ms < best_ms:
                best_ms, cur = ms, cand
    ctx.set_tuning(**cur)
    return {**cur, "ms_per_step": round(best_ms, 2)}, rows


def load_table(path=TABLE):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {"entries": {}}


def lookup(len1, len2, path=TABLE):
    """The table's setting for this batch shape, or None."""
    e = load_table(path)["entries"].get(shape_key(len1, len2))
    return {k: e["tuning"].get(k, DEFAULT[k]) for k in DEFAULT} if e else None


def main():
    import argparse
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import dcns_amd
    from dcns_amd import synth
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=4096)
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--write", metavar="PATH", nargs="?", const=TABLE, default=None,
                    help="merge the result into the table at PATH (default: the package's table)")
    args = ap.parse_args()
    pairs = synth.batch(npairs=args.pairs)
    l1 = [len(a) for a, _ in pairs]
    l2 = [len(b) for _, b in pairs]
    ctx = dcns_amd.Context(0)
    ctx.stage(pairs)
    best, rows = tune(ctx, reps=args.reps, log=lambda m: print(m, file=sys.stderr, flush=True))
    ctx.close()
    key = shape_key(l1, l2)
    out = {"shape": key, "best": best, "measured": rows}
    print(json.dumps(out))
    if args.write:
        table = load_table(args.write)
        table["entries"][key] = {"tuning": best, "measured": rows,
                                 "workload": f"dcns_amd.synth.batch(npairs={args.pairs}), cut1Gap defaults"}
        os.makedirs(os.path.dirname(os.path.abspath(args.write)), exist_ok=True)
        with open(args.write, "w") as f:
            json.dump(table, f, indent=1)


if __name__ == "__main__":
    main()

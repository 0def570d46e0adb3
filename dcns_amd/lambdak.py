"""Karlin-Altschul lambda and K for dCNS p-values: the permutation fit and a
learned surrogate.

The reference estimates lambda and K by permutation: ranSco scores random
fragment pairs with SmithWaterman (permutationLsqLambdaK.cpp:32-102) and an R
script fits y(x) = #(S >= x)/n ~ 1 - exp(-K N exp(-lambda x)) by
Levenberg-Marquardt (scripts/scriptToestimateKandLambda.r:4-15).  Here:

* ``permutation_fit`` -- the same method on the device: random fragment pairs
  of a given composition and lengths generated and scored on the GPU
  (dcns_sw_random_max), then the FP64 fit (dcns_fit_lambda_k) with
  N = len1 * len2.  This is the plain method, the fallback outside the
  surrogate's training hull, and the judge of the surrogate.
* ``Surrogate`` -- an MLP (dcns_mlp_*: 8 inputs -> 4 x 256 GELU -> ln lambda,
  ln K; BF16 matrix-core kernels, FP32 AdamW) trained on permutation fits of
  many (scoring, composition, lengths) points.  Inside its hull it answers a
  query with no DP cells at all; outside it falls back to permutation_fit.

Inputs of a query: match, mismatch, gap open, gap extend, long-gap open,
GC fraction, ln(reference fragment length), ln(query fragment length).  The
reference's permutation path scores with the one-piece SmithWaterman
(permutationLsqLambdaK.cpp:97), so the long-gap open does not change its
lambda or K; the input is carried so that two-piece settings are looked up
with their full scoring, and the training points sample it.
"""
import itertools
import json
import math
import os
import time

import numpy as np

from . import Mlp, fit_lambda_k_batch

GC_RANGE = (0.30, 0.70)
LEN_RANGE = (38, 2000)
OPEN2_RANGE = (-60.0, -30.0)
MATCH, MISMATCH, OPEN, EXT = (1, 3), (-4, -1), (-6, -2), (-3, -1)
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "lambdak_surrogate.npz")


def setting_grid():
    """The integer scoring settings of the sweep: match 1..3, mismatch -1..-4,
    gap open -2..-6, gap extend -1..-3 (180 settings)."""
    return [s for s in itertools.product(range(MATCH[0], MATCH[1] + 1), range(MISMATCH[1], MISMATCH[0] - 1, -1),
                                         range(OPEN[1], OPEN[0] - 1, -1), range(EXT[1], EXT[0] - 1, -1))]


def log_uniform(rng, lo, hi, n):
    return np.exp(rng.uniform(math.log(lo), math.log(hi + 1), n)).astype(np.int64).clip(lo, hi)


def permutation_scores(ctx, setting, gc, len1, len2, pairs, seed):
    """Max local scores of `pairs` random fragment pairs per point
    ([points x pairs], int32), generated and scored on the device."""
    gc = np.atleast_1d(np.asarray(gc, np.float64))
    l1 = np.broadcast_to(np.asarray(len1, np.int32), gc.shape)
    l2 = np.broadcast_to(np.asarray(len2, np.int32), gc.shape)
    with np.errstate(over="ignore"):  # uint64 arithmetic wraps on purpose
        seeds = (np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15) + np.arange(len(gc), dtype=np.uint64)
                 * np.uint64(1 << 32)) if np.isscalar(seed) else np.asarray(seed, np.uint64)
    return ctx.sw_random_max(tuple(int(v) for v in setting[:4]), gc, l1, l2, seeds, pairs)


def permutation_fit(ctx, setting, gc, len1, len2, pairs=10000, seed=1):
    """The plain method for one or many points: (lambda, K) arrays and the DP
    cells spent."""
    with np.errstate(over="ignore"):
        sc = permutation_scores(ctx, setting, gc, len1, len2, pairs, seed)
    l1 = np.broadcast_to(np.asarray(len1, np.float64), (sc.shape[0],))
    l2 = np.broadcast_to(np.asarray(len2, np.float64), (sc.shape[0],))
    fit = fit_lambda_k_batch(sc, l1 * l2)
    return fit[:, 1], fit[:, 0], float(np.sum(l1 * l2) * pairs)


def features(setting, open2, gc, len1, len2):
    """Network inputs [n x 8] for n queries."""
    gc = np.atleast_1d(np.asarray(gc, np.float64))
    n = len(gc)
    s = np.broadcast_to(np.asarray(setting, np.float64)[:4], (n, 4))
    return np.column_stack([s, np.broadcast_to(np.asarray(open2, np.float64), (n,)), gc,
                            np.log(np.broadcast_to(np.asarray(len1, np.float64), (n,))),
                            np.log(np.broadcast_to(np.asarray(len2, np.float64), (n,)))]).astype(np.float32)


def regime_check(ctx, settings, pairs=512, seed=99):
    """Settings whose random-pair maximum stays logarithmic in the search
    space at both composition extremes (the Karlin-Altschul regime): the mean
    maximum of 2000 x 2000 pairs is below 3 times that of 200 x 200 pairs
    scaled by ln N (a linear regime grows ~10x between the two)."""
    keep = []
    for s in settings:
        ok = True
        for g in GC_RANGE:
            small = permutation_scores(ctx, s, [g], 200, 200, pairs, seed).mean()
            large = permutation_scores(ctx, s, [g], 2000, 2000, pairs, seed + 1).mean()
            if large > 3.0 * max(small, 1.0) * math.log(4e6) / math.log(4e4):
                ok = False
                break
        if ok:
            keep.append(tuple(s))
    return keep


def training_set(ctx, settings, points_per_setting, pairs_per_point=1024, seed=2026, log=None):
    """Permutation fits at random (composition, lengths) points of each
    setting.  Returns X [n x 8], Y [n x 2] = (ln lambda, ln K), and the DP
    cells spent."""
    rng = np.random.default_rng(seed)
    Xs, Ys, cells = [], [], 0.0
    for si, s in enumerate(settings):
        t = time.perf_counter()
        n = points_per_setting
        gc = rng.uniform(*GC_RANGE, n)
        l1, l2 = log_uniform(rng, *LEN_RANGE, n), log_uniform(rng, *LEN_RANGE, n)
        order = np.argsort(-(l1 * l2), kind="stable")  # similar work per wave
        gc, l1, l2 = gc[order], l1[order], l2[order]
        open2 = rng.uniform(*OPEN2_RANGE, n)
        lam, K, c = permutation_fit(ctx, s, gc, l1, l2, pairs_per_point, seed=seed * 1000 + si)
        good = np.isfinite(lam) & np.isfinite(K) & (lam > 0) & (K > 0)
        Xs.append(features(s, open2, gc, l1, l2)[good])
        Ys.append(np.column_stack([np.log(lam), np.log(K)])[good].astype(np.float32))
        cells += c
        if log:
            log(f"setting {si + 1}/{len(settings)} {s}: {n} points, {c:.3e} cells, {time.perf_counter() - t:.2f} s")
    return np.concatenate(Xs), np.concatenate(Ys), cells


class Surrogate:
    """The trained network plus its hull: the scoring settings it was trained
    on, the composition and length ranges, and the long-gap open range."""

    def __init__(self, ctx, settings, seed=0):
        self.ctx = ctx
        self.settings = sorted({tuple(int(v) for v in s) for s in settings})
        self.net = Mlp(ctx, seed=seed)
        self.scaling = None

    # -- training ---------------------------------------------------------
    def fit(self, X, Y, steps=20000, batch=4096, lr=2e-3, weight_decay=1e-4, seed=0, val_frac=0.1, chunk=500,
            lambda_weight=1.0, log=None):
        """AdamW on the device with a cosine learning-rate decay.  The targets
        are noisy permutation fits, so a held-out fraction of the points
        (val_frac) is scored after every `chunk` steps and the parameters with
        the lowest held-out loss are kept: past that point the network starts
        to fit the noise of individual targets.

        lambda_weight scales the ln(lambda) term of the squared-error loss
        against ln(K)'s: its output is standardised by std / sqrt(weight).
        A fit's ln(K) scatters ~10x more than its ln(lambda) in standardised
        units, so at weight 1 the loss (and the held-out choice) is almost
        all ln(K) noise."""
        X = np.asarray(X, np.float32)
        Y = np.asarray(Y, np.float32)
        sd = X.std(0)
        sd[sd == 0] = 1
        ys = Y.std(0)
        ys[ys == 0] = 1
        ys[0] /= math.sqrt(lambda_weight)
        self.scaling = (X.mean(0), sd, Y.mean(0), ys)
        self.net.set_scaling(*self.scaling)
        perm = np.random.default_rng(seed).permutation(len(X))
        nv = int(len(X) * val_frac)
        Xv, Yv = X[perm[:nv]], Y[perm[:nv]]
        Xt, Yt = X[perm[nv:]], Y[perm[nv:]]
        losses, best, best_step, best_params = [], np.inf, 0, None
        t = time.perf_counter()
        for k in range(0, steps, chunk):
            n = min(chunk, steps - k)
            a = lr * 0.5 * (1 + math.cos(math.pi * (k + n / 2) / steps))
            losses.append(self.net.train(Xt, Yt, n, batch, a, weight_decay, seed + k))
            vl = self.net.loss_grad(Xv, Yv)[0] if nv else float(np.mean(losses[-1][-50:]))
            if vl < best:
                best, best_step, best_params = vl, k + n, self.net.get_params()
            if log:
                log(f"step {k + n}: train {float(np.mean(losses[-1][-50:])):.5f} held-out {vl:.5f} lr {a:.2e} "
                    f"({time.perf_counter() - t:.1f} s)")
        if best_params is not None:
            self.net.set_params(best_params)
        self.fit_info = {"best_step": best_step, "held_out_loss": float(best), "held_out_points": nv}
        return np.concatenate(losses) if losses else np.zeros(0, np.float32)

    # -- queries ----------------------------------------------------------
    def in_hull(self, X):
        X = np.asarray(X, np.float64)
        known = np.array([tuple(int(round(v)) for v in row[:4]) in self._setset() and
                          np.allclose(row[:4], np.round(row[:4])) for row in X], bool)
        gc_ok = (X[:, 5] >= GC_RANGE[0]) & (X[:, 5] <= GC_RANGE[1])
        len_ok = ((X[:, 6] >= math.log(LEN_RANGE[0]) - 1e-6) & (X[:, 6] <= math.log(LEN_RANGE[1]) + 1e-6) &
                  (X[:, 7] >= math.log(LEN_RANGE[0]) - 1e-6) & (X[:, 7] <= math.log(LEN_RANGE[1]) + 1e-6))
        o2_ok = (X[:, 4] >= OPEN2_RANGE[0]) & (X[:, 4] <= OPEN2_RANGE[1])
        return known & gc_ok & len_ok & o2_ok

    def _setset(self):
        if not hasattr(self, "_ss"):
            self._ss = set(self.settings)
        return self._ss

    def predict(self, X):
        """(lambda, K) from the network alone (no hull test)."""
        Y = self.net.predict(np.asarray(X, np.float32)).astype(np.float64)
        return np.exp(Y[:, 0]), np.exp(Y[:, 1])

    def lambda_k(self, X, fallback_pairs=10000, seed=7):
        """(lambda, K, from_network): the network inside the hull, the
        permutation fit outside it."""
        X = np.asarray(X, np.float64)
        inside = self.in_hull(X)
        lam = np.zeros(len(X))
        K = np.zeros(len(X))
        if inside.any():
            lam[inside], K[inside] = self.predict(X[inside])
        for i in np.nonzero(~inside)[0]:
            row = X[i]
            s = tuple(int(round(v)) for v in row[:4])
            l, k, _ = permutation_fit(self.ctx, s, [row[5]], int(round(math.exp(row[6]))),
                                      int(round(math.exp(row[7]))), fallback_pairs, seed + int(i))
            lam[i], K[i] = l[0], k[0]
        return lam, K, inside

    # -- persistence ------------------------------------------------------
    def save(self, path=DEFAULT_WEIGHTS, meta=None):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.savez_compressed(path, params=self.net.get_params(), in_mean=self.scaling[0], in_std=self.scaling[1],
                            out_mean=self.scaling[2], out_std=self.scaling[3],
                            settings=np.array(self.settings, np.int32), meta=json.dumps(meta or {}))

    @classmethod
    def load(cls, ctx, path=DEFAULT_WEIGHTS):
        z = np.load(path, allow_pickle=False)
        s = cls(ctx, [tuple(r) for r in z["settings"]])
        s.scaling = (z["in_mean"], z["in_std"], z["out_mean"], z["out_std"])
        s.net.set_scaling(*s.scaling)
        s.net.set_params(z["params"])
        s.meta = json.loads(str(z["meta"]))
        return s


def pvalue(lam, K, len1, len2, score):
    """E = K * L1 * L2 * exp(-lambda * S), p = 1 - exp(-E) (x_extend_seed,
    findSimilarFragmentsForPairedSequence4.cpp:282-283)."""
    e = K * len1 * len2 * np.exp(-lam * score)
    return -np.expm1(-e), e


def train_main(argv=None):
    """Trains the surrogate on the GPU and writes its weights:

        python -m dcns_amd.lambdak --points 200000 --out dcns_amd/data/lambdak_surrogate.npz
    """
    import argparse
    import dcns_amd
    ap = argparse.ArgumentParser()
    ap.add_argument("--points", type=int, default=200000)
    ap.add_argument("--pairs", type=int, default=2048, help="random pairs per training point")
    ap.add_argument("--steps", type=int, default=30000)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--lambda-weight", type=float, default=1.0, help="loss weight of ln(lambda) against ln(K)")
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--data", default="", help="optional .npz to save / reuse the training set")
    a = ap.parse_args(argv)

    def log(m):
        print(m, flush=True)

    ctx = dcns_amd.Context(0)
    t0 = time.perf_counter()
    if a.data and os.path.exists(a.data):
        z = np.load(a.data, allow_pickle=False)
        X, Y, settings, cells = z["X"], z["Y"], [tuple(r) for r in z["settings"]], float(z["cells"])
        log(f"training set {a.data}: {len(X)} points")
    else:
        grid = setting_grid()
        settings = regime_check(ctx, grid)
        log(f"{len(settings)} of {len(grid)} settings in the logarithmic regime ({time.perf_counter() - t0:.1f} s)")
        per = int(math.ceil(a.points / len(settings)))
        X, Y, cells = training_set(ctx, settings, per, a.pairs, a.seed, log)
        if a.data:
            np.savez_compressed(a.data, X=X, Y=Y, settings=np.array(settings, np.int32), cells=cells)
    t1 = time.perf_counter()
    log(f"{len(X)} training points, {cells:.3e} DP cells, {t1 - t0:.1f} s")
    sur = Surrogate(ctx, settings, seed=a.seed)
    losses = sur.fit(X, Y, steps=a.steps, batch=a.batch, lr=a.lr, seed=a.seed, lambda_weight=a.lambda_weight, log=log)
    t2 = time.perf_counter()
    sur.save(a.out, meta={"points": int(len(X)), "pairs_per_point": a.pairs, "training_cells": cells,
                          "steps": a.steps, "batch": a.batch, "lr": a.lr, "seed": a.seed,
                          "lambda_weight": a.lambda_weight,
                          "data_seconds": round(t1 - t0, 1), "train_seconds": round(t2 - t1, 1),
                          "final_loss": float(np.mean(losses[-200:])), **sur.fit_info})
    log(f"trained in {t2 - t1:.1f} s, weights -> {a.out}")
    ctx.close()


if __name__ == "__main__":
    train_main()


def evaluate(ctx, sur, settings, points_per_setting=100, judge_pairs=10000, cheap_pairs=1024, seed=31,
             matched_pairs=0, log=None):
    """The surrogate against the permutation fit ("judge", judge_pairs random
    pairs per point) at random (composition, lengths) points of each setting,
    beside a permutation fit at the training points' budget (cheap_pairs)
    and, when matched_pairs > 0, one at that many pairs (a fit of about the
    surrogate's accuracy, whose cost per query prices the surrogate's
    speedup).  All three fits draw independent pairs.

    Returns a dict of per-point arrays and summary figures: relative errors
    of lambda and K, the error of the E-value at the score where the judge's
    p-value is 0.1 (the significance threshold), DP cells per query of each
    method, and the wall-clock cost per query."""
    rng = np.random.default_rng(seed)
    rows = []
    t_judge = t_matched = 0.0
    cells_judge = 0.0
    cells_cheap = 0.0
    cells_matched = 0.0
    lm_all, km_all = [], []
    for si, s in enumerate(settings):
        n = points_per_setting
        gc = rng.uniform(*GC_RANGE, n)
        l1, l2 = log_uniform(rng, *LEN_RANGE, n), log_uniform(rng, *LEN_RANGE, n)
        order = np.argsort(-(l1 * l2), kind="stable")
        gc, l1, l2 = gc[order], l1[order], l2[order]
        open2 = rng.uniform(*OPEN2_RANGE, n)
        t = time.perf_counter()
        with np.errstate(over="ignore"):
            sj = permutation_scores(ctx, s, gc, l1, l2, judge_pairs, seed=seed * 7919 + si)
        N = l1.astype(np.float64) * l2
        fj = fit_lambda_k_batch(sj, N)
        t_judge += time.perf_counter() - t
        lj, kj, cj = fj[:, 1], fj[:, 0], float(np.sum(N) * judge_pairs)
        h = judge_pairs // 2  # the judge's own scatter: fits of its two halves
        f1 = fit_lambda_k_batch(np.ascontiguousarray(sj[:, :h]), N)
        f2 = fit_lambda_k_batch(np.ascontiguousarray(sj[:, h:2 * h]), N)
        lc, kc, cc = permutation_fit(ctx, s, gc, l1, l2, cheap_pairs, seed=seed * 104729 + si)
        if matched_pairs:
            t = time.perf_counter()
            lm, km, cm = permutation_fit(ctx, s, gc, l1, l2, matched_pairs, seed=seed * 15485863 + si)
            t_matched += time.perf_counter() - t
            cells_matched += cm
            lm_all.append(lm)
            km_all.append(km)
        cells_judge += cj
        cells_cheap += cc
        X = features(s, open2, gc, l1, l2)
        inside = sur.in_hull(X)
        ls, ks = sur.predict(X)
        for i in range(n):
            rows.append((si, gc[i], l1[i], l2[i], lj[i], kj[i], lc[i], kc[i], ls[i], ks[i],
                         f1[i, 1], f1[i, 0], f2[i, 1], f2[i, 0], bool(inside[i])))
        if log:
            log(f"judge {si + 1}/{len(settings)} {s}: {n} points, hull {int(inside.sum())}/{n}")
    r = np.array([x[:14] for x in rows], np.float64)
    inside = np.array([x[14] for x in rows], bool)
    lj, kj, lc, kc, ls, ks = r[:, 4], r[:, 5], r[:, 6], r[:, 7], r[:, 8], r[:, 9]
    # half-sample fits differ by ~2x the full judge's error (sqrt 2 for half the
    # pairs, sqrt 2 for the difference of two), so |h1 - h2| / 2 estimates it
    judge_noise_l = np.abs(r[:, 10] / r[:, 12] - 1) / 2
    judge_noise_k = np.abs(r[:, 11] / r[:, 13] - 1) / 2
    N = r[:, 2] * r[:, 3]
    # score at which the judge's p-value is 0.1 for a search space N
    s_thr = np.log(kj * N / -np.log(0.9)) / lj
    e_j = kj * N * np.exp(-lj * s_thr)

    def e_err(lam, K):
        return np.abs(K * N * np.exp(-lam * s_thr) / e_j - 1)

    def summ(a):
        a = a[np.isfinite(a)]
        return {"median": float(np.median(a)), "p90": float(np.quantile(a, 0.9)), "max": float(a.max())} \
            if len(a) else None

    res = {
        "points": int(len(r)), "in_hull": int(inside.sum()),
        "surrogate": {"lambda_rel_err": summ(np.abs(ls / lj - 1)[inside]), "K_rel_err": summ(np.abs(ks / kj - 1)[inside]),
                      "evalue_rel_err_at_p0.1": summ(e_err(ls, ks)[inside]),
                      "within_2pct_lambda": float(np.mean(np.abs(ls / lj - 1)[inside] <= 0.02)) if inside.any() else None,
                      "within_2pct_K": float(np.mean(np.abs(ks / kj - 1)[inside] <= 0.02)) if inside.any() else None,
                      "cells_per_query": 0},
        "permutation_cheap": {"pairs": cheap_pairs, "lambda_rel_err": summ(np.abs(lc / lj - 1)),
                              "K_rel_err": summ(np.abs(kc / kj - 1)), "evalue_rel_err_at_p0.1": summ(e_err(lc, kc)),
                              "cells_per_query": cells_cheap / len(r)},
        "judge": {"pairs": judge_pairs, "cells_per_query": cells_judge / len(r),
                  "seconds_per_query": t_judge / len(r), "own_lambda_rel_err": summ(judge_noise_l),
                  "own_K_rel_err": summ(judge_noise_k)},
        "rows": r, "inside": inside,
    }
    if matched_pairs:
        lm, km = np.concatenate(lm_all), np.concatenate(km_all)
        res["permutation_matched"] = {"pairs": matched_pairs, "lambda_rel_err": summ(np.abs(lm / lj - 1)),
                                      "K_rel_err": summ(np.abs(km / kj - 1)),
                                      "evalue_rel_err_at_p0.1": summ(e_err(lm, km)),
                                      "cells_per_query": cells_matched / len(r),
                                      "seconds_per_query": t_matched / len(r)}
    return res


def inference_rate(sur, batch=65536, reps=5, seed=5):
    """Surrogate queries per second at a batch of `batch` (device forward,
    host-to-device inputs and device-to-host outputs included)."""
    rng = np.random.default_rng(seed)
    s = np.array(sur.settings)[rng.integers(0, len(sur.settings), batch)]
    X = np.column_stack([s, rng.uniform(*OPEN2_RANGE, batch), rng.uniform(*GC_RANGE, batch),
                         np.log(log_uniform(rng, *LEN_RANGE, batch)),
                         np.log(log_uniform(rng, *LEN_RANGE, batch))]).astype(np.float32)
    sur.net.predict(X)
    t = time.perf_counter()
    for _ in range(reps):
        sur.net.predict(X)
    dt = (time.perf_counter() - t) / reps
    return batch / dt, dt

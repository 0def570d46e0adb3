"""The FP64 lambda/K least-squares fit (host code) against scipy's MINPACK
Levenberg-Marquardt on the reference's model (scripts/scriptToestimateKandLambda.r:4-15).
CPU only."""
import numpy as np
import pytest
from scipy.optimize import least_squares

import dcns_amd


def survival(scores):
    s = np.sort(np.asarray(scores))
    x = np.unique(s)
    y = (len(s) - np.searchsorted(s, x, side="left")) / len(s)
    return x.astype(float), y


def linearised(x, y, N):
    """Start point from ln(-ln(1-y)) = ln(K N) - lambda x (MINPACK diverges
    from the reference's K = 0.3, lambda = 0.2 on some of these data)."""
    ok = (y > 0) & (y < 1)
    slope, icpt = np.polyfit(x[ok], np.log(-np.log(1 - y[ok])), 1)
    return [np.exp(icpt) / N, -slope]


def gumbel_scores(rng, n, lam, K, N):
    # P(S >= x) = 1 - exp(-K N e^{-lam x}) for integer scores
    u = rng.random(n)
    x = np.log(K * N / -np.log(u)) / lam
    return np.floor(x).astype(np.int32)


@pytest.mark.parametrize("lam,K", [(0.38, 0.0067), (0.30, 0.0023), (0.6, 0.05)])
def test_fixed_space_fit_matches_minpack(lam, K):
    rng = np.random.default_rng(int(lam * 1000))
    N = 1e6
    scores = gumbel_scores(rng, 20000, lam, K, N)
    x, y = survival(scores)
    f = lambda p: y - (1 - np.exp(-p[0] * N * np.exp(-p[1] * x)))
    ref = least_squares(f, linearised(x, y, N), method="lm", xtol=1.49e-8, ftol=1.49e-8, max_nfev=20000)
    k, l, rss, it = dcns_amd.fit_lambda_k(scores, fixed_space=N)
    assert abs(l - ref.x[1]) / ref.x[1] < 1e-5 and abs(k - ref.x[0]) / ref.x[0] < 1e-4, (k, l, ref.x)
    assert abs(rss - np.sum(ref.fun ** 2)) <= 1e-9 + 1e-6 * rss
    # the estimate lands near the generating parameters
    assert abs(l - lam) / lam < 0.05


def test_variable_space_fit_matches_exact_mixture():
    rng = np.random.default_rng(5)
    n = 8000
    m = np.exp(rng.uniform(np.log(38), np.log(2000), n))
    q = np.exp(rng.uniform(np.log(38), np.log(2000), n))
    Ns = np.floor(m) * np.floor(q)
    lam, K = 0.35, 0.01
    u = rng.random(n)
    scores = np.floor(np.log(K * Ns / -np.log(u)) / lam).astype(np.int32)
    x, y = survival(scores)
    f = lambda p: y - (1 - np.exp(-p[0] * Ns[None, :] * np.exp(-p[1] * x[:, None]))).mean(axis=1)
    ref = least_squares(f, linearised(x, y, np.exp(np.log(Ns).mean())), method="lm", xtol=1.49e-8, ftol=1.49e-8,
                        max_nfev=20000)
    k, l, rss, it = dcns_amd.fit_lambda_k(scores, log_space=np.log(Ns))
    # binning the search spaces into 1024 log bins changes the optimum by far less than the sampling noise
    assert abs(l - ref.x[1]) / ref.x[1] < 2e-4 and abs(k - ref.x[0]) / ref.x[0] < 2e-3, (k, l, ref.x)
    assert abs(l - lam) / lam < 0.05


def test_fit_rejects_bad_input():
    with pytest.raises(dcns_amd.DcnsError):
        dcns_amd.fit_lambda_k(np.zeros(0, np.int32))


def test_calculate_lambda_closed_forms():
    """Ungapped lambda: +1/-1 on uniform composition solves 0.25 e^l + 0.75 e^-l = 1,
    i.e. l = ln 3; other scorings and compositions against a bracketing root
    finder on the same equation."""
    import dcns_amd
    from scipy.optimize import brentq
    uni = np.array([0, 1, 3, 4] * 250, np.int8)
    assert abs(dcns_amd.calculate_lambda(uni, 1, -1) - np.log(3)) < 1e-9
    rng = np.random.default_rng(3)
    for match, mism, gc in ((2, -3, 0.5), (1, -2, 0.35), (3, -4, 0.62), (2, -1, 0.45)):
        s = rng.choice([1, 3, 0, 4], 20000, p=[gc / 2, gc / 2, (1 - gc) / 2, (1 - gc) / 2]).astype(np.int8)
        s = np.concatenate([s, [2, 2, 2]]).astype(np.int8)  # N is not counted
        f = np.array([np.mean(s[s != 2] == c) for c in (0, 1, 3, 4)])
        F = lambda x: sum(f[i] * f[j] * np.exp((match if i == j else mism) * x) for i in range(4) for j in range(4)) - 1
        want = brentq(F, 1e-6, 20)
        assert abs(dcns_amd.calculate_lambda(s, match, mism) - want) < 1e-7

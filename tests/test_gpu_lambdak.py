"""The lambda/K surrogate against the permutation fit on the GPU: inside its
hull it beats a permutation fit of the same budget per point as its training
targets (at no DP cells per query), and outside the hull it falls back to the
permutation fit itself."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sur(gpu):
    from dcns_amd import lambdak as lk
    return lk.Surrogate.load(gpu)


def test_surrogate_beats_permutation_at_training_budget(gpu, sur):
    from dcns_amd import lambdak as lk
    rng = np.random.default_rng(11)
    settings = [(2, -3, -4, -2), (1, -2, -3, -1)]
    budget = int(sur.meta["pairs_per_point"])  # random pairs behind each training target
    errs_s, errs_c, errs_d, errs_ks, errs_kc = [], [], [], [], []
    for k, s in enumerate(settings):
        assert s in set(sur.settings)
        n = 12
        gc = rng.uniform(0.3, 0.7, n)
        l1, l2 = lk.log_uniform(rng, 100, 2000, n), lk.log_uniform(rng, 100, 2000, n)
        print("judge", s, flush=True)
        lj, kj, _ = lk.permutation_fit(gpu, s, gc, l1, l2, pairs=100000, seed=500 + k)
        lc, kc, _ = lk.permutation_fit(gpu, s, gc, l1, l2, pairs=budget, seed=600 + k)
        X = lk.features(s, -45, gc, l1, l2)
        ls, ks, inside = sur.lambda_k(X)
        assert inside.all()
        errs_s += list(np.abs(ls / lj - 1))
        errs_c += list(np.abs(lc / lj - 1))
        errs_d += list(np.abs(0.382291 / lj - 1))  # the reference's default constant for every setting
        errs_ks += list(np.abs(ks / kj - 1))
        errs_kc += list(np.abs(kc / kj - 1))
    ms, mc, md = np.median(errs_s), np.median(errs_c), np.median(errs_d)
    print("lambda median rel err: surrogate", ms, f"permutation({budget})", mc, "default constant", md)
    assert ms < 0.02 and ms < mc and ms < md
    assert np.median(errs_ks) < np.median(errs_kc)


def test_outside_hull_falls_back_to_permutation(gpu, sur):
    from dcns_amd import lambdak as lk
    cases = [((3, -1, -2, -1), 0.5, 500, 500),  # scoring not in the hull (linear regime)
             ((2, -3, -4, -2), 0.5, 5000, 800),  # fragment longer than the hull
             ((2, -3, -4, -2), 0.8, 500, 500)]  # composition outside the hull
    X = np.concatenate([lk.features(s, -45, [g], a, b) for s, g, a, b in cases])
    lam, K, inside = sur.lambda_k(X, fallback_pairs=2048, seed=9)
    assert not inside.any()
    for i, (s, g, a, b) in enumerate(cases):
        l2, k2, _ = lk.permutation_fit(gpu, s, [g], a, b, pairs=2048, seed=9 + i)
        assert lam[i] == l2[0] and K[i] == k2[0]


def test_hull_bounds(sur):
    from dcns_amd import lambdak as lk
    X = np.concatenate([lk.features((2, -3, -4, -2), -45, [0.3, 0.7], [38, 2000], [2000, 38]),
                        lk.features((2, -3, -4, -2), -45, [0.29], [500], [500]),
                        lk.features((2, -3, -4, -2), -45, [0.5], [37], [500]),
                        lk.features((2.5, -3, -4, -2), -45, [0.5], [500], [500])])
    assert list(sur.in_hull(X)) == [True, True, False, False, False]
    assert sur.meta["pairs_per_point"] == 8192 and sur.meta["lambda_weight"] == 10.0

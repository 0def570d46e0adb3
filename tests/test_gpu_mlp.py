"""The surrogate network's device kernels (BF16 matrix-core GEMMs, GELU
epilogues, backward, AdamW) against the CPU oracle (oracle/mlp_oracle.py).

The oracle's bf16 emulation rounds exactly where the kernels round, so the
remaining difference is FP32 accumulation order plus rare one-ulp bf16 flips
of values that land on a rounding boundary: tolerances are normwise relative
errors of 2e-3 (outputs) and 1e-2 (gradients).  A transposed operand, a wrong
MFMA lane mapping or a dropped bias gives errors of order 1."""
import numpy as np
import pytest

from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def _setup(gpu, seed, n):
    import dcns_amd
    rng = np.random.default_rng(seed)
    X = np.column_stack([rng.uniform(1, 3, n), rng.uniform(-4, -1, n), rng.uniform(-6, -2, n), rng.uniform(-3, -1, n),
                         rng.uniform(-60, -30, n), rng.uniform(0.3, 0.7, n), rng.uniform(3.6, 7.6, n),
                         rng.uniform(3.6, 7.6, n)]).astype(np.float32)
    Y = np.column_stack([np.sin(X[:, 0] + X[:, 5]) - 1.0, X[:, 6] * 0.1 - 3 + np.cos(X[:, 1])]).astype(np.float32)
    # scaling of the sampled ranges (not of the drawn rows: n may be 1)
    sc = (np.array([2, -2.5, -4, -2, -45, 0.5, 5.6, 5.6], np.float32),
          np.array([0.58, 0.87, 1.15, 0.58, 8.7, 0.115, 1.15, 1.15], np.float32),
          np.array([-1.2, -2.4], np.float32), np.array([0.6, 0.8], np.float32))
    m = dcns_amd.Mlp(gpu, seed=seed)
    m.set_scaling(*sc)
    p = mo.init_params(seed).astype(np.float32)
    m.set_params(p)
    return m, X, Y, sc, p.astype(np.float64)


def test_mlp_shapes_and_param_roundtrip(gpu):
    m, X, Y, sc, p = _setup(gpu, 1, 64)
    assert m.n_params == mo.N_PARAMS == 200194
    assert np.array_equal(m.get_params(), p.astype(np.float32))
    import dcns_amd
    m2 = dcns_amd.Mlp(gpu, seed=5)
    q = m2.get_params()
    # library init: U(-1/sqrt(fan_in), +1/sqrt(fan_in)) per layer
    for (W, b), fan in zip(mo.unflatten(q), mo.SIZES[:5]):
        bound = 1 / np.sqrt(fan)
        assert np.abs(W).max() <= bound and np.abs(b).max() <= bound and W.std() > bound / 3


@pytest.mark.parametrize("n", [1, 31, 1000])
def test_mlp_forward_matches_oracle(gpu, n):
    m, X, Y, sc, p = _setup(gpu, 2, n)
    dev = m.predict(X).astype(np.float64)
    y, _ = mo.forward(p, X, sc, bf16_emulation=True)
    ref = y * sc[3] + sc[2]
    assert _rel(dev, ref) < 2e-3, _rel(dev, ref)
    # the fp64 network without rounding is close too, but only to bf16 precision
    y64, _ = mo.forward(p, X, sc)
    assert _rel(dev, y64 * sc[3] + sc[2]) < 5e-2


def test_mlp_predict_crosses_batch_chunks(gpu):
    m, X, Y, sc, p = _setup(gpu, 3, 70000)
    dev = m.predict(X).astype(np.float64)
    rows = np.r_[0:4, 65530:65542, 69996:70000]
    y, _ = mo.forward(p, X[rows], sc, bf16_emulation=True)
    assert _rel(dev[rows], y * sc[3] + sc[2]) < 2e-3


@pytest.mark.parametrize("n", [7, 300])
def test_mlp_gradient_matches_oracle(gpu, n):
    m, X, Y, sc, p = _setup(gpu, 4, n)
    loss, g = m.loss_grad(X, Y)
    rl, rg = mo.loss_and_grad(p, X, Y, sc, bf16_emulation=True)
    assert abs(loss - rl) <= 2e-3 * rl
    offs = np.cumsum([0] + [mo.SIZES[i] * mo.SIZES[i + 1] + mo.SIZES[i + 1] for i in range(5)])
    for li in range(5):
        a, b = offs[li], offs[li + 1]
        nw = mo.SIZES[li] * mo.SIZES[li + 1]
        assert _rel(g[a:a + nw], rg[a:a + nw]) < 1e-2, ("W", li)
        assert _rel(g[a + nw:b], rg[a + nw:b]) < 1e-2, ("b", li)
    # and the exact fp64 gradient of the fp64 network, to bf16 precision
    _, g64 = mo.loss_and_grad(p, X, Y, sc)
    assert _rel(g, g64) < 5e-2


def test_mlp_adamw_step_matches_oracle(gpu):
    m, X, Y, sc, p = _setup(gpu, 5, 200)
    _, g = m.loss_grad(X, Y)
    m.train(X, Y, steps=1, batch=len(X), lr=2e-3, weight_decay=0.05)
    dev = m.get_params().astype(np.float64)
    ref, _, _ = mo.adamw_step(p, g.astype(np.float64), np.zeros_like(p), np.zeros_like(p), 1, 2e-3, 0.05)
    assert np.allclose(dev, ref, rtol=1e-5, atol=2e-7)


def test_mlp_training_reduces_loss(gpu):
    m, X, Y, sc, p = _setup(gpu, 6, 4096)
    before, _ = m.loss_grad(X, Y)
    losses = m.train(X, Y, steps=400, batch=512, lr=1e-3, weight_decay=1e-4, seed=1)
    after, _ = m.loss_grad(X, Y)
    assert np.all(np.isfinite(losses))
    assert after < 0.05 * before, (before, after)


@pytest.mark.parametrize("n", [1, 333, 70000])
def test_mlp_fused_inference_equals_layer_path(gpu, n):
    """predict runs the fused inference kernel, loss_grad the layer-by-layer
    GEMMs.  With identity scaling and targets set to predict's own outputs,
    the loss is exactly 0 only if both give the same bits on every row
    (70,000 rows span two predict chunks)."""
    import dcns_amd
    rng = np.random.default_rng(9 + n)
    m = dcns_amd.Mlp(gpu, seed=9)
    m.set_scaling(np.zeros(8, np.float32), np.ones(8, np.float32), np.zeros(2, np.float32), np.ones(2, np.float32))
    X = rng.normal(0, 1, (n, 8)).astype(np.float32)
    Y = m.predict(X)
    assert np.isfinite(Y).all() and Y.std() > 0 if n > 1 else np.isfinite(Y).all()
    loss, _ = m.loss_grad(X, Y)
    assert loss == 0.0, loss
    Y2 = Y.copy()
    Y2[n // 2, 1] += 1e-3  # one changed target: the loss sees it
    assert m.loss_grad(X, Y2)[0] > 0

"""Pins the CPU oracle of the lambda/K surrogate network (oracle/mlp_oracle.py):
float64 gradients against central finite differences and torch autograd, the
bf16 rounding against torch's bfloat16 conversion, and AdamW against
torch.optim.AdamW."""
import numpy as np
import pytest

from oracle import mlp_oracle as mo


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, 8)) * 2 + 1
    Y = rng.normal(size=(n, 2))
    scaling = (rng.normal(size=8), rng.uniform(0.5, 2, 8), rng.normal(size=2), rng.uniform(0.5, 2, 2))
    return X, Y, scaling


def test_param_count():
    assert mo.N_PARAMS == 200194
    p = mo.init_params(1)
    assert np.array_equal(mo.flatten(mo.unflatten(p)), p)


def test_gradient_vs_finite_differences():
    X, Y, sc = _data(5, 3)
    p = mo.init_params(4)
    _, g = mo.loss_and_grad(p, X, Y, sc)
    rng = np.random.default_rng(5)
    # parameters from every layer: weights and biases of layers 0, 2 and 4
    offs = np.cumsum([0] + [mo.SIZES[i] * mo.SIZES[i + 1] + mo.SIZES[i + 1] for i in range(5)])
    idx = np.concatenate([rng.integers(offs[i], offs[i + 1], 6) for i in range(5)])
    for k in idx:
        h = 1e-5
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (mo.loss_and_grad(pp, X, Y, sc)[0] - mo.loss_and_grad(pm, X, Y, sc)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 + 1e-5 * abs(fd), (k, fd, g[k])


def test_gradient_vs_torch_autograd():
    torch = pytest.importorskip("torch")
    X, Y, sc = _data(37, 6)
    p = mo.init_params(7)
    loss, g = mo.loss_and_grad(p, X, Y, sc)
    layers = [(torch.tensor(W, requires_grad=True), torch.tensor(b, requires_grad=True)) for W, b in mo.unflatten(p)]
    a = (torch.tensor(X) - torch.tensor(sc[0])) / torch.tensor(sc[1])
    for W, b in layers[:4]:
        a = torch.nn.functional.gelu(a @ W.T + b)
    y = a @ layers[4][0].T + layers[4][1]
    t = (torch.tensor(Y) - torch.tensor(sc[2])) / torch.tensor(sc[3])
    tl = ((y - t) ** 2).mean()
    tl.backward()
    tg = np.concatenate([np.concatenate([W.grad.numpy().ravel(), b.grad.numpy()]) for W, b in layers])
    assert abs(tl.item() - loss) < 1e-12
    assert np.allclose(g, tg, rtol=1e-9, atol=1e-14)


def test_bf16_rounding_vs_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(8)
    v = np.concatenate([rng.normal(size=10000) * 10 ** rng.uniform(-6, 6, 10000),
                        [0.0, -0.0, 1.0, 1.00390625, 1.005859375, -3.0e38, 65504.0]]).astype(np.float32)
    ours = mo.bf16(v)
    theirs = torch.tensor(v).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(ours, theirs)


def test_adamw_vs_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(9)
    layers = mo.unflatten(mo.init_params(10))
    flat = mo.flatten(layers)
    tp = [torch.nn.Parameter(torch.tensor(a)) for W, b in layers for a in (W, b)]
    decay = [t for i, t in enumerate(tp) if i % 2 == 0]
    nodecay = [t for i, t in enumerate(tp) if i % 2 == 1]
    opt = torch.optim.AdamW([{"params": decay, "weight_decay": 0.01}, {"params": nodecay, "weight_decay": 0.0}],
                            lr=3e-3, betas=(0.9, 0.999), eps=1e-8)
    m1 = np.zeros_like(flat)
    m2 = np.zeros_like(flat)
    for t in range(1, 4):
        g = rng.normal(size=flat.size)
        gl = mo.unflatten(g)
        for i, (W, b) in enumerate(gl):
            tp[2 * i].grad = torch.tensor(W)
            tp[2 * i + 1].grad = torch.tensor(b)
        opt.step()
        flat, m1, m2 = mo.adamw_step(flat, g, m1, m2, t, 3e-3, 0.01)
    ours = flat
    theirs = np.concatenate([t.detach().numpy().ravel() for t in tp])
    assert np.allclose(ours, theirs, rtol=1e-10, atol=1e-12)

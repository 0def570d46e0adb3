"""CPU oracle of the lambda/K surrogate network -- TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and bench.py's CPU baseline may import this
module; the library never calls it.  The network has no counterpart in the
reference (which obtains lambda and K by permutation only, ranSco in
controlLayer4.cpp:2088-2149 and permutationLsqLambdaK.cpp:32-102), so the
oracle states the mathematics directly:

    x0 = (X - mu_in) / sd_in                      (8 inputs)
    z_l = a_{l-1} W_l^T + b_l,  a_l = GELU(z_l)    l = 0..3, width 256
    y = a_3 W_4^T + b_4                            (2 outputs)
    loss = mean over rows and outputs of (y - (Y - mu_out) / sd_out)^2
    GELU(z) = z * Phi(z)  (exact, erf form)

forward()/backward() compute in float64.  With bf16_emulation=True they round to
bfloat16 (round to nearest even) exactly where the device kernels do: the
standardised input, the weights, each hidden activation, the output-layer
gradient and each pre-activation gradient that feeds a matrix product; the
bias gradients use the unrounded pre-activation gradients (the device sums
its FP32 copies).  What remains between oracle and device is FP32
accumulation order.

Pinning: backward() in float64 is checked against central finite
differences of forward() and against torch autograd (tests/test_oracle_mlp.py);
the bf16 rounding is checked against torch's bfloat16 conversion.

Parameters are a flat float array in the library's logical order: per layer
W [out x in] row-major, then b [out].
"""
import math

import numpy as np

SIZES = [8, 256, 256, 256, 256, 2]
N_PARAMS = sum(SIZES[i] * SIZES[i + 1] + SIZES[i + 1] for i in range(5))


def bf16(a):
    """Round float64/float32 values to the nearest bfloat16 (ties to even),
    returned as float64."""
    f = np.ascontiguousarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unflatten(flat):
    flat = np.asarray(flat, np.float64)
    out, o = [], 0
    for i in range(5):
        fi, fo = SIZES[i], SIZES[i + 1]
        W = flat[o:o + fo * fi].reshape(fo, fi)
        o += fo * fi
        b = flat[o:o + fo]
        o += fo
        out.append((W, b))
    assert o == len(flat) == N_PARAMS
    return out


def flatten(layers):
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in layers])


try:
    from scipy.special import erf as _erf
except ImportError:  # pragma: no cover
    _erf = np.vectorize(math.erf, otypes=[np.float64])


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def _round(a, on):
    return bf16(a) if on else a


def forward(flat, X, scaling=None, bf16_emulation=False):
    """Returns (standardised outputs [n x 2], cache for backward)."""
    layers = unflatten(flat)
    X = np.asarray(X, np.float64)
    mu_in, sd_in = (np.zeros(8), np.ones(8)) if scaling is None else (scaling[0], scaling[1])
    if bf16_emulation:  # the device standardises in FP32
        x32 = (X.astype(np.float32) - np.float32(mu_in)) * (np.float32(1.0) / np.asarray(sd_in, np.float32))
        a = bf16(x32)
    else:
        a = (X - mu_in) / sd_in
    acts, pres = [a], []
    for li in range(4):
        W, b = layers[li]
        z = a @ _round(W, bf16_emulation).T + b
        pres.append(z)
        a = _round(gelu(z), bf16_emulation)
        acts.append(a)
    W, b = layers[4]
    y = a @ _round(W, bf16_emulation).T + b
    return y, (layers, acts, pres)


def loss_and_grad(flat, X, Y, scaling=None, bf16_emulation=False):
    """Mean squared error of the standardised outputs and its gradient (flat)."""
    y, (layers, acts, pres) = forward(flat, X, scaling, bf16_emulation)
    Y = np.asarray(Y, np.float64)
    mu_out, sd_out = (np.zeros(2), np.ones(2)) if scaling is None else (scaling[2], scaling[3])
    t = (Y - mu_out) / sd_out
    n = len(X)
    d = y - t
    loss = float(np.mean(d * d))
    dnext = _round(2.0 * d / (n * 2), bf16_emulation)  # d loss / d y
    dnext_f = dnext
    grads = [None] * 5
    for li in range(4, -1, -1):
        W, _ = layers[li]
        dW = dnext.T @ acts[li]
        db = dnext_f.sum(axis=0)
        grads[li] = (dW, db)
        if li == 0:
            break
        dact = dnext @ _round(W, bf16_emulation)
        dnext_f = dact * gelu_grad(pres[li - 1])
        dnext = _round(dnext_f, bf16_emulation)
    return loss, flatten(grads)


def adamw_step(flat, grad, m1, m2, t, lr, weight_decay, b1=0.9, b2=0.999, eps=1e-8):
    """One AdamW step (decoupled decay on weights only, bias-corrected
    moments); returns (params, m1, m2)."""
    decay = np.concatenate([np.concatenate([np.ones(W.size), np.zeros(b.size)]) for W, b in unflatten(flat)])
    p = flat - lr * weight_decay * flat * decay
    m1 = b1 * m1 + (1 - b1) * grad
    m2 = b2 * m2 + (1 - b2) * grad * grad
    c1, c2 = 1 - b1 ** t, 1 - b2 ** t
    p = p - lr * (m1 / c1) / (np.sqrt(m2 / c2) + eps)
    return p, m1, m2


def init_params(seed=0):
    """PyTorch Linear-style init, U(-1/sqrt(in), 1/sqrt(in)) (any seeded draw
    serves the tests; the library draws its own)."""
    rng = np.random.default_rng(seed)
    layers = []
    for i in range(5):
        bound = 1.0 / math.sqrt(SIZES[i])
        layers.append((rng.uniform(-bound, bound, (SIZES[i + 1], SIZES[i])), rng.uniform(-bound, bound, SIZES[i + 1])))
    return flatten(layers)

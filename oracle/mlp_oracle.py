"""CPU checker of the surrogate network (test infrastructure only).

The library's multilayer perceptron restated in FP64 numpy: y = x W^T + b
per layer with GELU (exact Gaussian CDF) after every hidden layer, the loss
(1/B) sum_b [ (1/S) sum_k (y_bk - t_bk)^2 + w BCE(y_bS, label_b) ] and its
gradients by backpropagation, and the Adam update of Kingma & Ba (2015).

With emulate=True the BF16 roundings of the device path are applied where
it rounds: the layer inputs and weights fed to the matrix products, the
GELU outputs, and the output/hidden gradients fed to the backward products
(the bias gradients use the unrounded FP32 gradients).  The products
themselves are exact here (FP64) and FP32-accumulated on the device.

Pinned by finite differences of the loss (emulate=False) and by the Adam
closed form, in tests/test_mlp_oracle.py.
"""
import numpy as np
from scipy.special import erf

SQRT1_2 = 0.7071067811865476
INV_SQRT_2PI = 0.3989422804014327


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as
    float64."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    nan = ((u & 0x7F800000) == 0x7F800000) & ((u & 0x007FFFFF) != 0)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, (u >> 16) | 0x40, r).astype(np.uint32) << 16
    return r.view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z * SQRT1_2))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z * SQRT1_2)) + z * INV_SQRT_2PI * np.exp(-0.5 * z * z)


def split(flat, dims):
    """[(W [out][in], b [out]), ...] from the flat parameter vector."""
    out, o = [], 0
    flat = np.asarray(flat, np.float64)
    for a, b in zip(dims[:-1], dims[1:]):
        W = flat[o:o + a * b].reshape(b, a)
        o += a * b
        out.append((W, flat[o:o + b]))
        o += b
    return out


def forward(layers, x, emulate=False):
    """Output y [B][n_out] and the per-layer (input, pre-activation)."""
    rnd = bf16 if emulate else (lambda v: np.asarray(v, np.float64))
    h = rnd(x)
    cache = []
    for i, (W, b) in enumerate(layers):
        z = h @ rnd(W).T + b
        cache.append((h, z))
        if i + 1 < len(layers):
            h = rnd(gelu(z).astype(np.float32) if emulate else gelu(z))
        else:
            h = z
    return h, cache


def loss_and_grads(layers, x, target, label, label_weight=1.0,
                   emulate=False):
    """Loss and [(dW, db), ...]."""
    rnd = bf16 if emulate else (lambda v: np.asarray(v, np.float64))
    y, cache = forward(layers, x, emulate)
    B = y.shape[0]
    S = y.shape[1] - 1
    t = np.asarray(target, np.float64)
    lab = np.asarray(label, np.float64)
    d = y[:, :S] - t
    yl = y[:, S]
    bce = np.maximum(yl, 0) - yl * lab + np.log1p(np.exp(-np.abs(yl)))
    loss = float(((d * d).sum(axis=1) / S + label_weight * bce).sum() / B)
    dy = np.empty_like(y)
    dy[:, :S] = 2.0 * d / (S * B)
    dy[:, S] = label_weight * (1.0 / (1.0 + np.exp(-yl)) - lab) / B
    if emulate:
        dy = dy.astype(np.float32).astype(np.float64)
    grads = [None] * len(layers)
    dz = dy
    for i in range(len(layers) - 1, -1, -1):
        W, _ = layers[i]
        h, _ = cache[i]
        dzr = rnd(dz)
        grads[i] = (dzr.T @ h, dz.sum(axis=0))
        if i == 0:
            break
        da = dzr @ rnd(W)
        _, zp = cache[i - 1]
        dz = da * gelu_grad(zp)
        if emulate:
            dz = dz.astype(np.float32).astype(np.float64)
    return loss, grads


def flatten(grads):
    return np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in grads])


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (t is the step number after the update, from 1)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v

"""CPU FP64 reference of the warm-start network (test infrastructure).

The network: widths w[0] .. w[L], y = W_L a_{L-1} + b_L with
a_l = GELU(W_l a_{l-1} + b_l), GELU(x) = x Phi(x).  `emulate=True` rounds,
exactly where the device kernels do, the inputs of every matrix product to
BF16 (round to nearest even): the input batch, the weights, the GELU outputs
and the back-propagated d(pre) — so the device and this reference differ
only by FP32 vs FP64 accumulation.  `emulate=False` is the plain FP64
network, pinned by central finite differences."""
import numpy as np
from scipy.special import erf

SQRT_HALF = 0.7071067811865476


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.ascontiguousarray(np.asarray(x, np.float64).astype(np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    u = np.where(nan, 0x7FC0, u).astype(np.uint32) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + erf(x * SQRT_HALF))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x * SQRT_HALF)) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def unpack(params, widths):
    out, off = [], 0
    for l in range(1, len(widths)):
        nw = widths[l] * widths[l - 1]
        W = params[off:off + nw].reshape(widths[l], widths[l - 1])
        off += nw
        b = params[off:off + widths[l]]
        off += widths[l]
        out.append((W, b))
    return out


def forward(params, widths, X, emulate=True):
    r = bf16 if emulate else (lambda v: np.asarray(v, np.float64))
    layers = unpack(np.asarray(params, np.float64), widths)
    a = r(X)
    ins, pres = [], []
    for l, (W, b) in enumerate(layers, start=1):
        ins.append(a)
        pre = a @ r(W).T + b
        pres.append(pre)
        if l < len(layers):
            a = r(gelu(pre))
    return pres[-1], (ins, pres)


def backward(params, widths, cache, dY, emulate=True):
    r = bf16 if emulate else (lambda v: np.asarray(v, np.float64))
    layers = unpack(np.asarray(params, np.float64), widths)
    ins, pres = cache
    grads = []
    d = np.asarray(dY, np.float64)
    for l in range(len(layers), 0, -1):
        W, _ = layers[l - 1]
        db = d.sum(axis=0)
        dW = r(d).T @ ins[l - 1]
        grads.append((dW, db))
        if l > 1:
            d = (r(d) @ r(W)) * gelu_grad(pres[l - 2])
    flat = []
    for dW, db in reversed(grads):
        flat += [dW.ravel(), db]
    return np.concatenate(flat)


def init_params(widths, rng, scale=1.0):
    """He-style initialisation: W ~ N(0, scale^2 / fan_in), b = 0."""
    parts = []
    for l in range(1, len(widths)):
        parts.append(rng.standard_normal(widths[l] * widths[l - 1]) * scale / np.sqrt(widths[l - 1]))
        parts.append(np.zeros(widths[l]))
    return np.concatenate(parts)

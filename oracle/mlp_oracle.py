"""CPU float64 checker for the closure MLP (46 -> 64 -> 64 -> 64 -> 6, ReLU).

Test infrastructure only: the tests compare the device kernels with it; the
product never calls it.  The closure has no counterpart in the reference.

Two forms:
- exact(): plain float64 forward/backward, no rounding.  Its gradient is
  pinned by central finite differences in the CPU tests.
- mirrored(): the same algebra with operands rounded to BF16
  (round-to-nearest-even via float32) exactly where the device stores BF16:
  inputs, weights, hidden activations, the loss gradient dy and the
  back-propagated dz.  Sums run in float64, so it differs from the device
  only by the device's FP32 accumulation order.

Parameters use the canonical layout of include/channel_closure.h:
W1[46][64], b1, W2[64][64], b2, W3[64][64], b3, W4[64][6], b4.
"""
import numpy as np

IN, H, OUT = 46, 64, 6
SHAPES = [(IN, H), (H,), (H, H), (H,), (H, H), (H,), (H, OUT), (OUT,)]
NPARAM = sum(int(np.prod(s)) for s in SHAPES)


def unpack(p):
    out, o = [], 0
    for s in SHAPES:
        n = int(np.prod(s))
        out.append(np.asarray(p[o:o + n], dtype=np.float64).reshape(s))
        o += n
    return out


def pack(arrs):
    return np.concatenate([np.asarray(a, dtype=np.float64).ravel() for a in arrs])


def bf16(x):
    """Round to BF16 (nearest even) through float32, returned as float64."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float64).astype(np.float32))
    b = f.view(np.uint32).astype(np.uint64)
    b = ((b + (((b >> 16) & 1) + 0x7FFF)) >> 16) << 16
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def init_params(seed=0):
    """He-uniform weights, zero biases."""
    rng = np.random.default_rng(seed)
    arrs = []
    for s in SHAPES:
        if len(s) == 2:
            lim = np.sqrt(6.0 / s[0])
            arrs.append(rng.uniform(-lim, lim, size=s))
        else:
            arrs.append(np.zeros(s))
    return pack(arrs)


def _run(p, X, T, rnd):
    r = rnd if rnd is not None else (lambda a: a)
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    W1, W2, W3, W4 = r(W1), r(W2), r(W3), r(W4)
    x = r(np.asarray(X, dtype=np.float64))
    z1 = x @ W1 + b1
    h1 = r(np.maximum(z1, 0.0))
    h2 = r(np.maximum(h1 @ W2 + b2, 0.0))
    h3 = r(np.maximum(h2 @ W3 + b3, 0.0))
    y = h3 @ W4 + b4
    if T is None:
        return y, None, None
    n = x.shape[0]
    d = y - T
    loss = float(np.sum(d * d) / (6.0 * n))
    dy = r(2.0 * d / (6.0 * n))
    dz3 = r((dy @ W4.T) * (h3 > 0))
    dz2 = r((dz3 @ W3.T) * (h2 > 0))
    dz1 = r((dz2 @ W2.T) * (h1 > 0))
    g = pack([x.T @ dz1, dz1.sum(0), h1.T @ dz2, dz2.sum(0), h2.T @ dz3, dz3.sum(0), h3.T @ dy, dy.sum(0)])
    return y, loss, g


def exact(p, X, T=None):
    """(y, loss, grad) in plain float64."""
    return _run(p, X, T, None)


def mirrored(p, X, T=None):
    """(y, loss, grad) with the device's BF16 storage points."""
    return _run(p, X, T, bf16)

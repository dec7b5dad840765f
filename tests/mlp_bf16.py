"""A numpy model of the learned-guide kernels' rounding points: inputs, weights,
post-ReLU activations and every backward delta rounded to BF16 (round to nearest even),
products accumulated in float64 here (the kernels accumulate in FP32).  The FP64 oracle
says what the network computes; this model says how far BF16 alone moves it, so a test
can hold the GPU to a much tighter bar than "BF16 is within a few percent"."""
from __future__ import annotations

import numpy as np

LAYERS = [(78, 64), (64, 64), (64, 64), (64, 64), (64, 40)]


def bf16(a) -> np.ndarray:
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params: np.ndarray):
    out, off = [], 0
    for i, o in LAYERS:
        out.append((params[off:off + i * o].reshape(o, i).astype(np.float64),
                    params[off + i * o:off + i * o + o].astype(np.float64)))
        off += i * o + o
    return out


def loss_grad(params: np.ndarray, x: np.ndarray, dirs: np.ndarray, w: np.ndarray):
    """x: (n, 78) encoded inputs; returns (loss, gradient) of the weighted mean NLL."""
    layers = unpack(params)
    n = len(x)
    a = bf16(x)
    acts = [a]
    out = None
    for l, (W, b) in enumerate(layers):
        z = a @ bf16(W).T + b
        if l < 4:
            a = bf16(np.maximum(z, 0.0))
            acts.append(a)
        else:
            out = z
    o = out.reshape(n, 8, 5)
    d = dirs.astype(np.float64)
    v = o[..., :3]
    ln = np.sqrt((v ** 2).sum(-1) + 1e-12)
    mu = v / ln[..., None]
    sp = np.where(o[..., 3] > 20, o[..., 3], np.log1p(np.exp(np.minimum(o[..., 3], 20))))
    kap = np.minimum(0.1 + sp, 1000.0)
    e = o[..., 4]
    pi = np.exp(e - e.max(1, keepdims=True))
    pi /= pi.sum(1, keepdims=True)
    c = (mu * d[:, None, :]).sum(-1)
    f = kap * np.exp(kap * (c - 1)) / (2 * np.pi * (1 - np.exp(-2 * kap)))
    q = (pi * f).sum(1)
    wn = w.astype(np.float64) / w.sum(dtype=np.float64)
    loss = float((-wn * np.log(q + 1e-20)).sum())
    dq = -wn / (q + 1e-20)
    g = np.zeros((n, 8, 5))
    g[..., 4] = dq[:, None] * pi * (f - q[:, None])
    e2 = np.exp(-2 * kap)
    dlogf = 1 / kap - 2 * e2 / (1 - e2) + (c - 1)
    g[..., 3] = np.where(0.1 + sp > 1000, 0.0, dq[:, None] * pi * f * dlogf / (1 + np.exp(-o[..., 3])))
    s = (dq[:, None] * pi * f * kap)[..., None] * d[:, None, :]
    g[..., :3] = (s - mu * (mu * s).sum(-1, keepdims=True)) / ln[..., None]
    D = bf16(g.reshape(n, 40))
    grads = [None] * 5
    for l in range(4, -1, -1):
        W, _ = layers[l]
        grads[l] = (D.T @ acts[l], D.sum(0))
        if l == 0:
            break
        D = bf16((D @ bf16(W)) * (acts[l] > 0))
    return loss, np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in grads])

"""Writes tests/golden/mlp_golden.npz: the learned-guide network evaluated by torch
autograd in float64 (an implementation independent of the oracle) on a small batch:
parameters, inputs, directions, weights, loss, gradient, and the parameters after one
Adam step (lr 1e-3, betas 0.9/0.999, eps 1e-8).  Run: python tests/golden/make_mlp_golden.py
"""
import math
from pathlib import Path

import numpy as np
import torch

IN, H, OUT, K = 78, 64, 40, 8
SHAPES = [(H, IN), (H, H), (H, H), (H, H), (OUT, H)]


def make(seed=0, B=16):
    g = np.random.default_rng(seed)
    params = []
    for o, i in SHAPES:
        params.append(g.normal(scale=1.0 / math.sqrt(i), size=(o, i)))
        params.append(g.normal(scale=0.1, size=o))
    flat = np.concatenate([p.ravel() for p in params])
    x = g.uniform(-1, 1, size=(B, IN))
    d = g.normal(size=(B, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    w = g.uniform(0.1, 3.0, size=B)
    return flat, x, d.astype(np.float32), w.astype(np.float32)


def torch_loss(flat, x, d, w):
    P = torch.tensor(flat, dtype=torch.float64, requires_grad=True)
    X = torch.tensor(x, dtype=torch.float64)
    D = torch.tensor(d.astype(np.float64))
    Wt = torch.tensor(w.astype(np.float64))
    off = 0
    h = X
    for l, (o, i) in enumerate(SHAPES):
        Wm = P[off:off + o * i].view(o, i); off += o * i
        b = P[off:off + o]; off += o
        h = h @ Wm.T + b
        if l < 4:
            h = torch.relu(h)
    out = h.view(-1, K, 5)
    v = out[..., :3]
    mu = v / torch.sqrt((v * v).sum(-1, keepdim=True) + 1e-12)
    kap = 0.1 + torch.nn.functional.softplus(out[..., 3])
    pi = torch.softmax(out[..., 4], dim=-1)
    c = (mu * D[:, None, :]).sum(-1)
    f = kap * torch.exp(kap * (c - 1)) / (2 * math.pi * (1 - torch.exp(-2 * kap)))
    q = (pi * f).sum(-1)
    wn = Wt / Wt.mean() / len(Wt)
    loss = -(wn * torch.log(q + 1e-20)).sum()
    loss.backward()
    return loss.item(), P.grad.numpy().copy()


def main():
    flat, x, d, w = make()
    loss, grad = torch_loss(flat, x, d, w)
    m = 0.1 * grad
    v = 0.01 * grad * grad
    m1 = 0.9 * m + 0.1 * grad
    v1 = 0.999 * v + 0.001 * grad * grad
    after = flat - 1e-3 * (m1 / (1 - 0.9 ** 3)) / (np.sqrt(v1 / (1 - 0.999 ** 3)) + 1e-8)
    out = Path(__file__).with_name("mlp_golden.npz")
    idx = np.random.default_rng(1).choice(len(flat), size=2000, replace=False)
    # adam state before the step is (0.1 g, 0.01 g^2) at t = 3; only a 2000-entry subset of
    # the gradient and of the updated parameters is kept, plus the gradient's full norm
    np.savez_compressed(out, params=flat, x=x, dirs=d, w=w, loss=np.float64(loss), idx=idx, grad_sub=grad[idx],
                        grad_norm=np.linalg.norm(grad), adam_after_sub=after[idx])
    print(f"wrote {out}: loss {loss:.6f}")


if __name__ == "__main__":
    main()

"""Writes tests/golden/mlp_golden.npz: the learned-guide network evaluated by torch
autograd in float64 (an implementation independent of the oracle) on a small batch:
parameters, inputs, directions, weights, BSDF densities, loss, gradient (with the selection
probability's blend term, and without it), and the parameters after one Adam step (lr 1e-3,
betas 0.9/0.999, eps 1e-8).  Run: python tests/golden/make_mlp_golden.py
"""
import math
from pathlib import Path

import numpy as np
import torch

IN, H, OUT, K = 78, 64, 41, 8
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
    pb = g.uniform(0.0, 1.0, size=B)
    pb[::5] = 0.0  # directions the BSDF cannot sample
    ps = g.uniform(0.05, 2.0, size=B)  # densities the directions were drawn from
    return flat, x, d.astype(np.float32), w.astype(np.float32), np.stack([pb, ps], 1).astype(np.float32)


def torch_loss(flat, x, d, w, pdfs=None):
    P = torch.tensor(flat, dtype=torch.float64, requires_grad=True)
    X = torch.tensor(x, dtype=torch.float64)
    D = torch.tensor(d.astype(np.float64))
    Wt = torch.tensor(w.astype(np.float64))
    off = 0
    h = X
    for l, (o, i) in enumerate(SHAPES):
        Wm = P[off:off + o * i].view(o, i); off += o * i
        b = P[off:off + o]; off += o
        if l == 4:
            s_sel = h.detach() @ Wm[K * 5] + b[K * 5]  # the selection output reads h4 held constant
        h = h @ Wm.T + b
        if l < 4:
            h = torch.relu(h)
    out = h[:, :K * 5].reshape(-1, K, 5)
    v = out[..., :3]
    mu = v / torch.sqrt((v * v).sum(-1, keepdim=True) + 1e-12)
    kap = 0.1 + torch.nn.functional.softplus(out[..., 3])
    pi = torch.softmax(out[..., 4], dim=-1)
    c = (mu * D[:, None, :]).sum(-1)
    f = kap * torch.exp(kap * (c - 1)) / (2 * math.pi * (1 - torch.exp(-2 * kap)))
    q = (pi * f).sum(-1)
    wn = Wt / Wt.mean() / len(Wt)
    loss = -(wn * torch.log(q + 1e-20)).sum()
    if pdfs is not None:  # one-sample MIS blend's relative second moment; trains the selection output only
        pb = torch.tensor(pdfs[:, 0].astype(np.float64))
        ps = torch.tensor(pdfs[:, 1].astype(np.float64))
        alpha = 0.05 + 0.9 * torch.sigmoid(s_sel)
        qb = alpha * q.detach() + (1 - alpha) * pb + 1e-20
        loss = loss + len(Wt) * (wn * wn * ps / qb).sum()
    loss.backward()
    return loss.item(), P.grad.numpy().copy()


def main():
    flat, x, d, w, pdfs = make()
    loss, grad = torch_loss(flat, x, d, w, pdfs)
    loss_nopb, grad_nopb = torch_loss(flat, x, d, w)
    m = 0.1 * grad
    v = 0.01 * grad * grad
    m1 = 0.9 * m + 0.1 * grad
    v1 = 0.999 * v + 0.001 * grad * grad
    after = flat - 1e-3 * (m1 / (1 - 0.9 ** 3)) / (np.sqrt(v1 / (1 - 0.999 ** 3)) + 1e-8)
    out = Path(__file__).with_name("mlp_golden.npz")
    idx = np.random.default_rng(1).choice(len(flat), size=2000, replace=False)
    w4 = sum(o * i + o for o, i in SHAPES[:4])  # the selection output's weight row and every output bias
    idx = np.unique(np.concatenate([idx, w4 + 40 * H + np.arange(H), np.arange(len(flat) - OUT, len(flat))]))
    # adam state before the step is (0.1 g, 0.01 g^2) at t = 3; only a 2000-entry subset of
    # the gradient and of the updated parameters is kept, plus the gradient's full norm
    np.savez_compressed(out, params=flat, x=x, dirs=d, w=w, pdfs=pdfs, loss=np.float64(loss), idx=idx,
                        grad_sub=grad[idx], grad_norm=np.linalg.norm(grad), adam_after_sub=after[idx],
                        loss_nopb=np.float64(loss_nopb), grad_sub_nopb=grad_nopb[idx],
                        grad_norm_nopb=np.linalg.norm(grad_nopb))
    print(f"wrote {out}: loss {loss:.6f}")


if __name__ == "__main__":
    main()

"""The guides at their configured sizes.

* vMF grid EM step: a 32^3 x 8-lobe grid over the indoor room's bounds and 2,097,152 samples
  (the size the guided path tracer feeds it at 1080p), sufficient statistics and fitted lobes
  against the oracle's E- and M-step (`orc_vmm_estep`, `orc_vmm_mstep`) with the bars of the
  small-grid tests.
* MLP: loss and gradient on a batch of 262,144 samples (the configured training batch) against
  the FP64 oracle (`orc_mlp_loss_grad`, run in chunks on the host threads and recombined by the
  chunks' weight sums, since the loss is the weight-normalised mean), with the bars of the
  batch cases in tests/test_gpu_mlp.py; and the full batch equal to the weighted combination of
  its two halves, which checks the kernels' cross-block reduction at this size.
"""
import concurrent.futures as cf
import os

import numpy as np
import pytest

from oracle import oracle as O
from research_amd import _lib as L
from research_amd import core
from tests.mlp_bf16 import bf16
from tests.test_gpu_mlp import BMAX, BMIN, _encode, _rel, _samples
from tests.test_oracle import _vmf_draw

pytestmark = pytest.mark.gpu


def test_vmm_em_step_at_32cubed_by_8_with_2m_samples(ctx):
    rng = np.random.default_rng(21)
    res, K, n = 32, 8, 2_097_152
    lo, hi = (0.0, 0.0, 0.0), (10.0, 4.0, 10.0)
    pos = (rng.uniform(0, 1, (n, 3)) * np.array(hi)).astype(np.float32)
    # three directional components whose mix depends on the position (light from a window side)
    comp = np.where(pos[:, 0] < 5.0, 0, 1) + (rng.random(n) < 0.25)
    dirs = np.empty((n, 3), np.float32)
    for c, (mu, k) in enumerate((([0.0, 1.0, 0.0], 12.0), ([1.0, 0.3, 0.0], 4.0), ([0.0, 0.0, -1.0], 30.0))):
        sel = comp == c
        dirs[sel] = _vmf_draw(mu, k, int(sel.sum()), rng)
    w = rng.uniform(0.1, 2.0, n).astype(np.float32)
    w[rng.random(n) < 0.1] = 0.0
    prm = L.GridParams(L.GRID_VMM, 0, 0, 0, 0, 512.0, res, K, L.f3(lo), L.f3(hi), 0.0, 0.0)
    g = core.Grid(ctx, prm)
    lobes0, _ = g.download()
    L.call("rrh_grid_em_step", ctx.h, g.h, L.ptr(pos), L.ptr(dirs), L.ptr(w), n)
    lobes_gpu, stats_gpu = g.download()
    st = np.zeros((res ** 3, K, 4), np.float64)
    lob = lobes0.copy()
    O.lib().orc_vmm_estep(O.p(lob), K, O.p(pos), O.p(dirs), O.p(w), n, O.fa(lo), O.fa(hi), res, O.p(st))
    assert (st[..., 0] > 0).any(axis=1).mean() > 0.99     # nearly every cell received samples
    assert np.allclose(stats_gpu, st, rtol=1e-4, atol=1e-6)
    O.lib().orc_vmm_mstep(O.p(st), res ** 3, K, 512.0, 0.0, O.p(lob))
    assert np.allclose(lobes_gpu, lob, rtol=2e-4, atol=2e-4)


def _oracle_loss_grad_chunked(wts, x, dirs, w, chunks=16):
    """orc_mlp_loss_grad over chunks in host threads; each chunk is normalised by its own
    weight sum, so the whole batch is the chunks' weighted mean"""
    lib = O.lib()
    P = np.ascontiguousarray(wts.astype(np.float64))
    bounds = np.linspace(0, len(w), chunks + 1).astype(int)

    def one(k):
        a, b = bounds[k], bounds[k + 1]
        g = np.zeros(core.Mlp.NPARAM)
        xs, ds, ws = (np.ascontiguousarray(v[a:b]) for v in (x, dirs, w))
        loss = lib.orc_mlp_loss_grad(O.p(P), O.p(xs), O.p(ds), O.p(ws), int(b - a), O.p(g))
        return float(ws.astype(np.float64).sum()), loss, g

    with cf.ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        parts = list(ex.map(one, range(chunks)))
    W = sum(p[0] for p in parts)
    return sum(p[0] * p[1] for p in parts) / W, sum(p[0] * p[2] for p in parts) / W


def test_mlp_loss_grad_at_the_configured_batch(ctx):
    n = 262_144
    rng = np.random.default_rng(7)
    m = core.Mlp(ctx, BMIN, BMAX, seed=3)
    wts = m.get_weights()
    wts[-40:] = np.tile(np.float32([0.3, -0.2, 0.5, 2.0, 0.0]), 8) + rng.normal(0, 0.3, 40).astype(np.float32)
    wts = bf16(wts).astype(np.float32)   # BF16-exact, so the oracle and the kernels share the network
    m.set_weights(wts)
    pos, nrm, wo, dirs, w = _samples(n, rng)
    loss, g = m.loss_grad(pos, nrm, wo, dirs, w)
    # halves: the full batch is their weight-normalised combination
    h = n // 2
    l1, g1 = m.loss_grad(pos[:h], nrm[:h], wo[:h], dirs[:h], w[:h])
    l2, g2 = m.loss_grad(pos[h:], nrm[h:], wo[h:], dirs[h:], w[h:])
    W1, W2 = float(w[:h].astype(np.float64).sum()), float(w[h:].astype(np.float64).sum())
    assert abs(loss - (W1 * l1 + W2 * l2) / (W1 + W2)) <= 1e-4 * abs(loss)
    assert _rel(g, (W1 * g1 + W2 * g2) / (W1 + W2)) <= 1e-4
    # against the FP64 oracle on the whole batch
    x = _encode(pos, nrm, wo)
    lref, gref = _oracle_loss_grad_chunked(wts, x, dirs, w)
    cos = float(g @ gref / (np.linalg.norm(g) * np.linalg.norm(gref)))
    print(f"n={n}: loss {loss} oracle {lref}; gradient cosine {cos:.6f}, relative L2 {_rel(g, gref):.4f}")
    assert abs(loss - lref) <= 1e-2 * abs(lref), (loss, lref)
    assert cos >= 0.999 and _rel(g, gref) <= 3e-2, (cos, _rel(g, gref))
    assert m.state()[0] == 0  # no update

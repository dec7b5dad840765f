"""Pins for the FP64 learned-guide oracle: torch-autograd golden fixture (an
independent implementation), central finite differences, mixture-pdf normalisation,
and the Adam update against its closed form."""
import math
from pathlib import Path

import numpy as np

from oracle import oracle as O

lib = O.lib()
GOLD = Path(__file__).parent / "golden" / "mlp_golden.npz"


def _loss_grad(params, x, d, w):
    g = np.zeros_like(params)
    loss = lib.orc_mlp_loss_grad(O.p(params), O.p(x), O.p(d), O.p(w), len(w), O.p(g))
    return loss, g


def test_param_count():
    assert lib.orc_mlp_param_count() == 20136  # 78x64 + 3x64x64 + 64x40 weights + 296 biases


def test_matches_torch_autograd_fixture():
    z = np.load(GOLD)
    params = np.ascontiguousarray(z["params"])
    loss, g = _loss_grad(params, np.ascontiguousarray(z["x"]), np.ascontiguousarray(z["dirs"]),
                         np.ascontiguousarray(z["w"]))
    assert abs(loss - float(z["loss"])) < 1e-10
    assert np.allclose(g[z["idx"]], z["grad_sub"], rtol=1e-8, atol=1e-12)
    assert abs(np.linalg.norm(g) - float(z["grad_norm"])) < 1e-9 * float(z["grad_norm"])
    m, v = 0.1 * g, 0.01 * g * g
    p2 = params.copy()
    lib.orc_adam(O.p(p2), O.p(g), O.p(m), O.p(v), len(p2), 3, 1e-3, 0.9, 0.999, 1e-8)
    assert np.allclose(p2[z["idx"]], z["adam_after_sub"], rtol=1e-12, atol=1e-14)


def test_gradient_matches_finite_differences():
    z = np.load(GOLD)
    params = np.ascontiguousarray(z["params"]).copy()
    x, d, w = (np.ascontiguousarray(z[k])[:4] for k in ("x", "dirs", "w"))
    _, g = _loss_grad(params, x, d, w)
    rng = np.random.default_rng(2)
    idx = np.concatenate([rng.choice(20136, 150, replace=False), np.arange(20136 - 40, 20136)])
    h = 1e-6
    for i in idx:
        p = params[i]
        params[i] = p + h
        lp, _ = _loss_grad(params, x, d, w)
        params[i] = p - h
        lm, _ = _loss_grad(params, x, d, w)
        params[i] = p
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_head_mixture_is_a_density():
    z = np.load(GOLD)
    params = np.ascontiguousarray(z["params"])
    acts, out, lobes = np.zeros(256), np.zeros(40), np.zeros(40)
    x0 = np.ascontiguousarray(z["x"][0])
    lib.orc_mlp_forward(O.p(params), O.p(x0), O.p(acts), O.p(out))
    lib.orc_mlp_head(O.p(out), O.p(lobes))
    L = lobes.reshape(8, 5)
    assert abs(L[:, 4].sum() - 1) < 1e-12 and (L[:, 3] >= 0.1).all()
    assert np.allclose(np.linalg.norm(L[:, :3], axis=1), 1, atol=1e-9)
    n = 200000
    i = np.arange(n) + 0.5
    ph = np.arccos(1 - 2 * i / n)
    th = np.pi * (1 + 5 ** 0.5) * i
    dirs = np.stack([np.cos(th) * np.sin(ph), np.sin(th) * np.sin(ph), np.cos(ph)], 1).astype(np.float32)
    vals = np.array([lib.orc_mlp_pdf(O.p(lobes), O.fa(dd)) for dd in dirs[::10]])
    assert abs(vals.mean() * 4 * math.pi - 1) < 5e-3


def test_encoding_frequencies():
    x = np.zeros(78)
    lib.orc_mlp_encode(O.fa([0.25, 0.5, 1.0]), O.fa([0, 0, 0]), O.fa([1, 1, 2]), O.fa([0, 0, 1]),
                       O.fa([1, 0, 0]), O.p(x))
    u = [0.25, 0.5, 0.5]
    for c in range(3):
        for j in range(12):
            assert abs(x[c * 24 + 2 * j] - math.sin(2 ** j * math.pi * u[c])) < 1e-12
            assert abs(x[c * 24 + 2 * j + 1] - math.cos(2 ** j * math.pi * u[c])) < 1e-12
    assert x[72:75].tolist() == [0, 0, 1] and x[75:78].tolist() == [1, 0, 0]

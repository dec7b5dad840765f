"""Pins for the FP64 learned-guide oracle: torch-autograd golden fixture (an
independent implementation), central finite differences, mixture-pdf normalisation,
the selection probability's closed form and the optimum of its blend term, and the Adam
update against its closed form."""
import math
from pathlib import Path

import numpy as np

from oracle import oracle as O

lib = O.lib()
GOLD = Path(__file__).parent / "golden" / "mlp_golden.npz"


NP = 20201
W4 = 78 * 64 + 64 + 3 * (64 * 64 + 64)  # offset of the output layer's weights


def _loss_grad(params, x, d, w, pb=None):
    g = np.zeros_like(params)
    loss = lib.orc_mlp_loss_grad(O.p(params), O.p(x), O.p(d), O.p(w), None if pb is None else O.p(pb), len(w),
                                 O.p(g))
    return loss, g


def test_param_count():
    assert lib.orc_mlp_param_count() == NP  # 78x64 + 3x64x64 + 64x41 weights + 297 biases


def test_matches_torch_autograd_fixture():
    z = np.load(GOLD)
    params = np.ascontiguousarray(z["params"])
    args = [np.ascontiguousarray(z[k]) for k in ("x", "dirs", "w")]
    for pb, sfx in ((np.ascontiguousarray(z["pdfs"]), ""), (None, "_nopb")):
        loss, g = _loss_grad(params, *args, pb)
        assert abs(loss - float(z["loss" + sfx])) < 1e-10
        assert np.allclose(g[z["idx"]], z["grad_sub" + sfx], rtol=1e-8, atol=1e-12)
        assert abs(np.linalg.norm(g) - float(z["grad_norm" + sfx])) < 1e-9 * float(z["grad_norm" + sfx])
    g = _loss_grad(params, *args, np.ascontiguousarray(z["pdfs"]))[1]
    m, v = 0.1 * g, 0.01 * g * g
    p2 = params.copy()
    lib.orc_adam(O.p(p2), O.p(g), O.p(m), O.p(v), len(p2), 3, 1e-3, 0.9, 0.999, 1e-8)
    assert np.allclose(p2[z["idx"]], z["adam_after_sub"], rtol=1e-12, atol=1e-14)


def test_gradient_matches_finite_differences():
    """Without densities the loss is the mixture NLL, so central differences pin the whole
    gradient; with them the blend term holds q and the last hidden layer constant, which
    differences of the loss cannot see, so there only the selection output's row and bias are
    compared (the other parameters are pinned by the next test)."""
    z = np.load(GOLD)
    params = np.ascontiguousarray(z["params"]).copy()
    x, d, w, pb = (np.ascontiguousarray(z[k])[:5] for k in ("x", "dirs", "w", "pdfs"))
    rng = np.random.default_rng(2)
    cases = [(None, np.concatenate([rng.choice(NP, 150, replace=False), np.arange(NP - 41, NP)])),
             (pb, np.concatenate([W4 + 40 * 64 + rng.choice(64, 16, replace=False), [NP - 1]]))]
    h = 1e-6
    for pbc, idx in cases:
        _, g = _loss_grad(params, x, d, w, pbc)
        for i in idx:
            p = params[i]
            params[i] = p + h
            lp, _ = _loss_grad(params, x, d, w, pbc)
            params[i] = p - h
            lm, _ = _loss_grad(params, x, d, w, pbc)
            params[i] = p
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_blend_term_trains_the_selection_output_only():
    """The blend term holds the mixture and the last hidden layer constant: every parameter but
    the selection output's row and bias gets the same gradient with and without the densities,
    and that row none without them."""
    z = np.load(GOLD)
    params = np.ascontiguousarray(z["params"])
    x, d, w, pb = (np.ascontiguousarray(z[k]) for k in ("x", "dirs", "w", "pdfs"))
    _, g0 = _loss_grad(params, x, d, w)
    _, g1 = _loss_grad(params, x, d, w, pb)
    sel_row = np.r_[W4 + 40 * 64:W4 + 41 * 64, NP - 1]
    rest = np.setdiff1d(np.arange(NP), sel_row)
    assert np.array_equal(g0[rest], g1[rest])
    assert not g0[sel_row].any() and np.abs(g1[sel_row]).max() > 0


def test_selection_closed_form_and_blend_optimum():
    """alpha = 0.05 + 0.9 sigmoid(s); with one alpha for every sample the blend term
    n sum_s wn_s^2 ps_s / (alpha q_s + (1 - alpha) pb_s) is convex in alpha, and its derivative in
    s vanishes where sum_s wn_s^2 ps_s (q_s - pb_s) / (alpha q_s + (1 - alpha) pb_s)^2 = 0 (found
    here by bisection, independent of the oracle's gradient)."""
    out = np.zeros(41)
    for s_, a in ((0.0, 0.5), (math.log(3.0), 0.05 + 0.9 * 0.75), (-50.0, 0.05), (50.0, 0.95)):
        out[40] = s_
        assert abs(lib.orc_mlp_select(O.p(out)) - a) < 1e-12
    z = np.load(GOLD)
    params = np.ascontiguousarray(z["params"]).copy()
    x, d, w, pb = (np.ascontiguousarray(z[k]) for k in ("x", "dirs", "w", "pdfs"))
    # every sample gets the same selection output: zero the selection row, steer with its bias
    params[W4 + 40 * 64:W4 + 41 * 64] = 0.0
    q = np.zeros(len(w))
    acts, o, lob = np.zeros(256), np.zeros(41), np.zeros(40)
    for i in range(len(w)):
        lib.orc_mlp_forward(O.p(params), O.p(np.ascontiguousarray(x[i])), O.p(acts), O.p(o))
        lib.orc_mlp_head(O.p(o), O.p(lob))
        q[i] = lib.orc_mlp_pdf(O.p(lob), O.fa(d[i]))
    wn = w / w.sum()
    qd = np.float64(q)
    pbd, psd = pb[:, 0].astype(np.float64), pb[:, 1].astype(np.float64)

    def dalpha(a):
        return float((wn * wn * psd * (qd - pbd) / (a * qd + (1 - a) * pbd) ** 2).sum())

    lo, hi = 0.05, 0.95
    if dalpha(lo) * dalpha(hi) < 0:
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            lo, hi = (mid, hi) if dalpha(mid) * dalpha(lo) > 0 else (lo, mid)
        a_opt = 0.5 * (lo + hi)
        params[NP - 1] = math.log((a_opt - 0.05) / (0.95 - a_opt))  # sigmoid^-1 of (a - 0.05) / 0.9
        g = _loss_grad(params, x, d, w, pb)[1]
        assert _loss_grad(params, x, d, w, None)[1][NP - 1] == 0.0
        params[NP - 1] = 0.0
        g_half = _loss_grad(params, x, d, w, pb)[1]
        # the blend term's derivative vanishes at the optimum (against its size at alpha = 0.5)
        assert abs(g[NP - 1]) < 1e-6 * abs(g_half[NP - 1]), (g[NP - 1], g_half[NP - 1])
    else:  # the optimum sits on a bound: the derivative pushes towards it
        params[NP - 1] = 0.0
        g = _loss_grad(params, x, d, w, pb)[1]
        assert g[NP - 1] * dalpha(0.5) < 0


def test_head_mixture_is_a_density():
    z = np.load(GOLD)
    params = np.ascontiguousarray(z["params"])
    acts, out, lobes = np.zeros(256), np.zeros(41), np.zeros(40)
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

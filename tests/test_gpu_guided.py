"""Guided path tracing (8-lobe EM grid and learned guide): the one-sample MIS estimator
must stay unbiased whatever the guide has learned, so the guided mean image has to agree
with the unguided one within their Monte Carlo error.  The guides are trained online
from the tracer's own recorded vertices."""
import ctypes as C

import numpy as np
import pytest

from research_amd import _lib as L
from research_amd import core

pytestmark = pytest.mark.gpu

W = H = 32
BOUNCES = 5


def _scene(ctx):
    sa = core.generate_scene(L.GEN_CORNELL_RANDOM, 1, 1.0)
    gs = core.Scene(ctx, sa)
    gs.set_camera(core.camera_look_at(sa.eye, sa.center, sa.fov, W, H))
    return gs


def _render(ctx, gs, img, seed, mode=L.GUIDE_NONE, alpha=0.0, record=0, grid=None, mlp=None):
    p = L.PTParams(BOUNCES, seed, 3, mode, alpha, record)
    L.call("rrh_pt_render_guided", ctx.h, gs.h, C.byref(p), grid.h if grid else None, mlp.h if mlp else None, img.h)
    return img.download().reshape(H, W, 4)[..., :3].astype(np.float64)


def _frames(ctx, gs, n, seed0, **kw):
    img = core.Image(ctx, W, H)
    blocks = []
    for k in range(n):
        r = _render(ctx, gs, img, seed0 + k, **kw)
        assert np.isfinite(r).all()
        blocks.append(r.reshape(4, 8, 4, 8, 3).mean(axis=(1, 3, 4)).ravel())  # 16 block means
    return np.array(blocks)


def _assert_same_mean(a, b, what):
    ma, mb = a.mean(0), b.mean(0)
    se = np.sqrt(a.var(0, ddof=1) / len(a) + b.var(0, ddof=1) / len(b))
    z = np.abs(ma - mb) / np.maximum(se, 1e-12)
    assert z.max() < 5.0, (what, z.max(), ma, mb)
    tot_a, tot_b = a.mean(1), b.mean(1)
    zt = abs(tot_a.mean() - tot_b.mean()) / np.sqrt(tot_a.var(ddof=1) / len(a) + tot_b.var(ddof=1) / len(b))
    assert zt < 5.0, (what, zt)


@pytest.fixture(scope="module")
def unguided(ctx):
    gs = _scene(ctx)
    return gs, _frames(ctx, gs, 256, 1000)


def test_vmm_guided_render_is_unbiased(ctx, unguided):
    gs, ref = unguided
    info = gs.info()
    prm = L.GridParams(L.GRID_VMM, 0, 0, 0, 0, 512.0, 4, 8, L.f3(info["bounds_min"]), L.f3(info["bounds_max"]),
                       0.0, 0.0)
    g = core.Grid(ctx, prm)
    img = core.Image(ctx, W, H)
    for k in range(8):
        _render(ctx, gs, img, 50 + k, mode=L.GUIDE_VMM, alpha=0.5 if k else 0.0, record=20000, grid=g)
        L.call("rrh_grid_em_from_vertices", ctx.h, g.h)
    lobes, _ = g.download()
    assert np.isfinite(lobes).all() and (lobes[..., 3] > 0).any()
    got = _frames(ctx, gs, 256, 5000, mode=L.GUIDE_VMM, alpha=0.5, grid=g)
    _assert_same_mean(got, ref, "vmm")


def test_mlp_guided_render_is_unbiased_and_training_lowers_the_loss(ctx, unguided):
    gs, ref = unguided
    info = gs.info()
    m = core.Mlp(ctx, info["bounds_min"], info["bounds_max"], lr=2e-3, seed=1)
    img = core.Image(ctx, W, H)
    losses = []
    for k in range(60):
        _render(ctx, gs, img, 70 + k, mode=L.GUIDE_MLP, alpha=0.5 if k >= 20 else 0.0, record=4096, mlp=m)
        losses.append(m.train_from_vertices(4096))
    assert np.isfinite(losses).all()
    assert np.mean(losses[-10:]) < np.mean(losses[:5]), (losses[:5], losses[-10:])
    assert m.state()[0] >= 55
    got = _frames(ctx, gs, 256, 9000, mode=L.GUIDE_MLP, alpha=0.5, mlp=m)
    _assert_same_mean(got, ref, "mlp")


def test_guide_training_is_deterministic(ctx):
    """The recorded vertices are compacted in record order and the EM step sums per cell
    without atomics, so the same frames train the same guides: two fresh contexts (the vertex
    thinning hashes a per-context render stamp), each with a fresh VMM grid and MLP (same
    seed), fed the same guided frames, end bit-identical."""
    grids, mlps = [], []
    for _ in range(2):
        c2 = core.Context(0)
        gs = _scene(c2)
        info = gs.info()
        img = core.Image(c2, W, H)
        g = core.Grid(c2, L.GridParams(L.GRID_VMM, 0, 0, 0, 0, 512.0, 8, 8, L.f3(info["bounds_min"]),
                                       L.f3(info["bounds_max"]), 0.9, 0.0))
        m = core.Mlp(c2, info["bounds_min"], info["bounds_max"], lr=3e-3, seed=5)
        for k in range(4):
            _render(c2, gs, img, 700 + k, mode=L.GUIDE_VMM, alpha=0.5 if k else 0.0, record=2000, grid=g)
            L.call("rrh_grid_em_from_vertices", c2.h, g.h)
            _render(c2, gs, img, 800 + k, mode=L.GUIDE_MLP, alpha=0.5 if k else 0.0, record=1024, mlp=m)
            m.train_from_vertices(1024)
        grids.append(g.download())
        mlps.append(m.get_weights())
        for o in (m, g, img, gs):
            o.close()
        c2.close()
    assert np.array_equal(grids[0][0].view(np.uint32), grids[1][0].view(np.uint32))
    assert np.array_equal(grids[0][1].view(np.uint64), grids[1][1].view(np.uint64))
    assert np.array_equal(mlps[0].view(np.uint32), mlps[1].view(np.uint32))
    assert grids[0][1].any()  # the grids did see samples


def _noise(ctx, gs, w, h, n, seed0, bounces, **kw):
    """Per-sample noise: mean over pixels of the variance over n frames / (mean^2 + 0.01), and
    the mean image."""
    img = core.Image(ctx, w, h)
    acc, acc2 = np.zeros((h, w, 3)), np.zeros((h, w, 3))
    for k in range(n):
        p = L.PTParams(bounces, seed0 + k, 3, kw.get("mode", L.GUIDE_NONE), kw.get("alpha", 0.0), 0)
        L.call("rrh_pt_render_guided", ctx.h, gs.h, C.byref(p), kw["grid"].h if kw.get("grid") else None,
               kw["mlp"].h if kw.get("mlp") else None, img.h)
        r = img.download()[..., :3].astype(np.float64)
        acc += r
        acc2 += r * r
    m = acc / n
    var = (acc2 - n * m * m) / (n - 1)
    return float(np.mean(var / (m * m + 0.01))), m


def test_trained_guides_lower_the_noise_of_the_window_lit_room(ctx):
    """What the trained guides gain over the tracer without them: in the indoor room lit only
    through a small window, one-sample MIS with the online-EM vMF grid or the trained MLP gives
    clearly less variance per sample than BSDF sampling alone (and the same mean).  Both guides
    train online for 160 frames at 256x192 (about 0.4 M path vertices per frame for the grid, one
    65,536-vertex Adam step per frame at lr 3e-3 for the MLP); with the 64x48 evaluation extent
    alone they see too few paths and a poorly fitted guide adds variance.  The MLP needs the
    steps: over initialisation seeds 1-6 its noise was 0.85-1.09 after 48 frames, 0.60-0.85 after
    96 and 0.48-0.57 after 160 (unguided 1.098)."""
    W2, H2, B = 64, 48, 8
    sa = core.generate_scene(L.GEN_INDOOR, 1, 0.25)
    gs = core.Scene(ctx, sa)
    info = gs.info()
    grid = core.Grid(ctx, L.GridParams(L.GRID_VMM, 0, 0, 0, 0, 512.0, 32, 8, L.f3(info["bounds_min"]),
                                       L.f3(info["bounds_max"]), 0.95, 0.0))
    mlp = core.Mlp(ctx, info["bounds_min"], info["bounds_max"], lr=3e-3, seed=1)
    # online training from the tracer's own vertices, at a larger extent (more paths per frame)
    TW, TH = 256, 192
    gs.set_camera(core.camera_look_at(sa.eye, sa.center, sa.fov, TW, TH))
    timg = core.Image(ctx, TW, TH)
    for k in range(160):
        a = 0.5 if k else 0.0
        p = L.PTParams(B, 100 + k, 3, L.GUIDE_VMM, a, 400000)
        L.call("rrh_pt_render_guided", ctx.h, gs.h, C.byref(p), grid.h, None, timg.h)
        L.call("rrh_grid_em_from_vertices", ctx.h, grid.h)
        p = L.PTParams(B, 300 + k, 3, L.GUIDE_MLP, a, 65536)
        L.call("rrh_pt_render_guided", ctx.h, gs.h, C.byref(p), None, mlp.h, timg.h)
        mlp.train_from_vertices(65536)
    gs.set_camera(core.camera_look_at(sa.eye, sa.center, sa.fov, W2, H2))
    n = 256
    v_none, m_none = _noise(ctx, gs, W2, H2, n, 10000, B)
    v_vmm, m_vmm = _noise(ctx, gs, W2, H2, n, 20000, B, mode=L.GUIDE_VMM, alpha=0.5, grid=grid)
    v_mlp, m_mlp = _noise(ctx, gs, W2, H2, n, 30000, B, mode=L.GUIDE_MLP, alpha=0.5, mlp=mlp)
    print(f"per-sample noise: unguided {v_none:.3f}, vMF grid {v_vmm:.3f}, MLP {v_mlp:.3f}")
    assert v_vmm < 0.85 * v_none and v_mlp < 0.7 * v_none
    for m in (m_vmm, m_mlp):  # the same image (within noise): guiding changes the variance, not the mean
        assert abs(m.mean() / m_none.mean() - 1) < 0.1


def test_mlp_guided_frame_equal_through_the_traversal_epilogue_and_the_input_pass(ctx):
    """The network's per-slot inputs come from the traversal kernels' epilogue (paired launches:
    k_trace_pair; unpaired: k_intersect) or, where a traversal variant does not write them, from
    the separate input pass (k_mlp_prep, taken here with fetch counting).  Both write them through
    one function, so the guided frame, its ray counts and its recorded training set are the same
    bit for bit whichever path ran."""
    gs = _scene(ctx)
    info = gs.info()
    m = core.Mlp(ctx, info["bounds_min"], info["bounds_max"], seed=3)
    img = core.Image(ctx, W, H)

    def frame(flags, pairs):
        L.call("rrh_ctx_set_option", ctx.h, L.OPT_TRACE_PAIRS, pairs)
        try:
            p = L.PTParams(BOUNCES, 4242, 3, L.GUIDE_MLP, 0.5, 4096, flags)
            L.call("rrh_pt_render_guided", ctx.h, gs.h, C.byref(p), None, m.h, img.h)
            rc = L.RayCounts()
            L.call("rrh_pt_ray_counts", ctx.h, C.byref(rc))
            t = core.Mlp(ctx, info["bounds_min"], info["bounds_max"], seed=7)  # one step on the recorded set
            loss = t.train_from_vertices(4096)
            return img.download().copy(), (rc.primary, rc.extension, rc.shadow), (loss, t.get_weights().tobytes())
        finally:
            L.call("rrh_ctx_set_option", ctx.h, L.OPT_TRACE_PAIRS, L.TRACE_PAIRS_DEFAULT)

    ref = frame(L.PT_COUNT_FETCHES, 1)  # fetch counting: unpaired, counting kernels, input pass
    for pairs in (1, 0):
        got = frame(0, pairs)
        assert np.array_equal(got[0].view(np.uint32), ref[0].view(np.uint32)), pairs
        assert got[1] == ref[1], (pairs, got[1], ref[1])
        assert got[2] == ref[2], pairs

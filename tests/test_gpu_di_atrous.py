"""GPU ReSTIR DI and A-trous against the oracle."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from research_amd import _lib as L
from research_amd import core
from tests.helpers import cornell_sun_mr, ocam, scene_arrays

pytestmark = pytest.mark.gpu


def _bits_equal_frac(a, b):
    a = np.ascontiguousarray(a).reshape(len(a), -1).view(np.uint32)
    b = np.ascontiguousarray(b).reshape(len(b), -1).view(np.uint32)
    return np.all(a == b, axis=1).mean()


@pytest.mark.parametrize("kind,scale,M", [(L.GEN_CORNELL_REF, 1.0, 16), (L.GEN_CORNELL_RANDOM, 1.0, 32),
                                          (L.GEN_HEIGHTFIELD, 0.1, 32), (cornell_sun_mr, 1.0, 16)])
def test_restir_di_three_stages_match_oracle(ctx, kind, scale, M):
    sa = scene_arrays(kind, 2, scale)
    gs = core.Scene(ctx, sa)
    os_ = O.Scene.from_arrays(sa)
    W, H = 64, 40
    cam = core.camera_look_at(sa.eye, sa.center, sa.fov, W, H)
    gs.set_camera(cam)
    oc = ocam(cam)
    im = lambda fmt=L.FMT_RGBA32F: core.Image(ctx, W, H, fmt)
    gp, gn, out = im(), im(), im()
    res = [im(), im()]
    uv = [im(L.FMT_R32U), im(L.FMT_R32U)]
    o_prev = None
    for frame in range(3):
        cur, prev = frame & 1, 1 - (frame & 1)
        prm = L.DIParams(frame, M, 1, 1, 5, 30)
        oprm = O.DIParams(frame, M, 1, 1, 5, 30)
        L.call("rrh_di_initial", ctx.h, gs.h, C.byref(prm), gp.h, gn.h, res[cur].h, uv[cur].h)
        ogp, ogn, ors, ouv = os_.di_initial(oc, oprm)
        assert _bits_equal_frac(gp.download().reshape(-1, 4), ogp) == 1.0
        assert _bits_equal_frac(res[cur].download().reshape(-1, 4), ors) == 1.0
        if frame > 0:
            # temporal on the GPU's own previous reservoirs, oracle on the same inputs
            rp, rpuv = res[prev].download().reshape(-1, 4), uv[prev].download().reshape(-1)
            rc0, rcuv0 = res[cur].download().reshape(-1, 4), uv[cur].download().reshape(-1)
            L.call("rrh_di_temporal", ctx.h, gs.h, C.byref(prm), gp.h, gn.h, res[cur].h, uv[cur].h, res[prev].h,
                   uv[prev].h)
            g_gp, g_gn = gp.download().reshape(-1, 4), gn.download().reshape(-1, 4)
            orc, orcuv = os_.di_temporal(oc, oprm, g_gp, g_gn, rc0, rcuv0, rp, rpuv)
            assert _bits_equal_frac(res[cur].download().reshape(-1, 4), orc) == 1.0
            assert (uv[cur].download().reshape(-1) == orcuv).mean() == 1.0
        L.call("rrh_di_spatial_shade", ctx.h, gs.h, C.byref(prm), gp.h, gn.h, res[cur].h, uv[cur].h, out.h)
        g_out = out.download().reshape(-1, 4)
        o_out = os_.di_spatial_shade(oc, oprm, gp.download().reshape(-1, 4), gn.download().reshape(-1, 4),
                                     res[cur].download().reshape(-1, 4), uv[cur].download().reshape(-1))
        assert _bits_equal_frac(g_out, o_out) == 1.0
        assert np.isfinite(g_out).all()


def _ref_inputs(H, W, seed):
    rng = np.random.default_rng(seed)
    rad = (rng.random((H, W, 4)) * np.array([3, 2, 1, 1])).astype(np.float32)
    pos = np.zeros((H, W, 4), np.float32)
    yy, xx = np.mgrid[0:H, 0:W]
    pos[..., 0] = xx * 0.02
    pos[..., 1] = np.where(xx > W // 2, 0.3, 0.0)
    pos[..., 2] = yy * 0.02
    pos[..., 3] = 1.0
    pos[rng.random((H, W)) < 0.05, 3] = -1.0  # misses
    nrm = np.zeros((H, W, 4), np.float32)
    nrm[..., 1] = 1.0
    nrm[xx > W // 2] = [0.0, 0.6, 0.8, 0.0]
    return rad, pos, nrm


def test_atrous_reference_mode_matches_oracle(ctx):
    H, W = 72, 136
    rad, pos, nrm = _ref_inputs(H, W, 0)
    imgs = [core.Image(ctx, W, H) for _ in range(5)]
    for im, a in zip(imgs[:3], (rad, pos, nrm)):
        im.upload(a)
    for iters in (1, 5):
        prm = L.AtrousParams(iters, L.ATROUS_REFERENCE, 0.25, 0.2, 1.0)
        slot = C.c_int()
        L.call("rrh_atrous", ctx.h, C.byref(prm), imgs[0].h, imgs[1].h, imgs[2].h, imgs[3].h, imgs[4].h,
               C.byref(slot))
        assert slot.value == (iters - 1) & 1
        g = imgs[3 + slot.value].download()
        o = rad
        for k in range(iters):
            o = O.atrous_ref(o, pos, nrm, 1 << k)
        assert np.abs(g - o).max() <= 1e-3
        assert _bits_equal_frac(g.reshape(-1, 4), o.reshape(-1, 4)) == 1.0


def test_atrous_config_mode_matches_oracle(ctx):
    H, W = 64, 120
    rad, pos, nrm = _ref_inputs(H, W, 1)
    rad16 = rad.astype(np.float16).view(np.uint16).copy()
    rad16[..., 3] = np.float16(0.05).view(np.uint16)
    gb = np.zeros((H, W, 4), np.float32)
    gb[..., :3] = nrm[..., :3]
    gb[..., 3] = np.where(pos[..., 3] < 0, 0.0, 2.0 + pos[..., 2])
    gb16 = gb.astype(np.float16).view(np.uint16).copy()
    i_in, i_gb, t0, t1 = (core.Image(ctx, W, H, L.FMT_RGBA16F) for _ in range(4))
    i_in.upload(rad16)
    i_gb.upload(gb16)
    prm = L.AtrousParams(5, L.ATROUS_CONFIG, 1.0, 128.0, 4.0)
    slot = C.c_int()
    L.call("rrh_atrous", ctx.h, C.byref(prm), i_in.h, i_gb.h, None, t0.h, t1.h, C.byref(slot))
    g = (t0, t1)[slot.value].download()
    o = rad16
    for k in range(5):
        o = O.atrous_cfg(o, gb16, 1 << k)
    gf = g.view(np.float16).astype(np.float32)
    of = o.view(np.float16).astype(np.float32)
    assert np.abs(gf - of).max() <= 1e-3 * max(1.0, np.abs(of).max())
    assert (g == o).all(axis=2).mean() == 1.0


def test_atrous_fast_math_within_tolerance_of_the_exact_filter(ctx):
    """RRH_ATROUS_FAST_MATH (hardware exp / log / reciprocal / rsqrt in the weights) against the
    exact oracle: per pixel within 1e-3 relative (absolute below 1), plus one fp16 step of the
    output format in config mode, at 1, 3 and 5 iterations (strides up to 16)."""
    H, W = 72, 136
    rad, pos, nrm = _ref_inputs(H, W, 2)
    imgs = [core.Image(ctx, W, H) for _ in range(5)]
    for im, a in zip(imgs[:3], (rad, pos, nrm)):
        im.upload(a)
    for iters in (1, 3, 5):
        prm = L.AtrousParams(iters, L.ATROUS_REFERENCE, 0.25, 0.2, 1.0, L.ATROUS_FAST_MATH)
        slot = C.c_int()
        L.call("rrh_atrous", ctx.h, C.byref(prm), imgs[0].h, imgs[1].h, imgs[2].h, imgs[3].h, imgs[4].h,
               C.byref(slot))
        g = imgs[3 + slot.value].download()
        o = rad
        for k in range(iters):
            o = O.atrous_ref(o, pos, nrm, 1 << k)
        assert np.isfinite(g).all()
        assert (np.abs(g - o) <= 1e-3 * np.maximum(1.0, np.abs(o))).all()

    rad16 = rad.astype(np.float16).view(np.uint16).copy()
    rad16[..., 3] = np.float16(0.05).view(np.uint16)
    gb = np.zeros((H, W, 4), np.float32)
    gb[..., :3] = nrm[..., :3]
    gb[..., 3] = np.where(pos[..., 3] < 0, 0.0, 2.0 + pos[..., 2])
    gb16 = gb.astype(np.float16).view(np.uint16).copy()
    i_in, i_gb, t0, t1 = (core.Image(ctx, W, H, L.FMT_RGBA16F) for _ in range(4))
    i_in.upload(rad16)
    i_gb.upload(gb16)
    for iters in (1, 5):
        prm = L.AtrousParams(iters, L.ATROUS_CONFIG, 1.0, 128.0, 4.0, L.ATROUS_FAST_MATH)
        slot = C.c_int()
        L.call("rrh_atrous", ctx.h, C.byref(prm), i_in.h, i_gb.h, None, t0.h, t1.h, C.byref(slot))
        g = (t0, t1)[slot.value].download()
        o = rad16
        for k in range(iters):
            o = O.atrous_cfg(o, gb16, 1 << k)
        of16 = o.view(np.float16)
        gf = g.view(np.float16).astype(np.float32)
        of = of16.astype(np.float32)
        step = np.spacing(np.abs(of16)).astype(np.float32)
        assert np.isfinite(gf).all()
        assert (np.abs(gf - of) <= 1e-3 * np.maximum(1.0, np.abs(of)) + step).all()


def test_atrous_rejects_unknown_flags(ctx):
    im = [core.Image(ctx, 8, 8) for _ in range(5)]
    prm = L.AtrousParams(1, L.ATROUS_REFERENCE, 0.25, 0.2, 1.0, 2)
    with pytest.raises(L.RRHError, match="unknown flags"):
        L.call("rrh_atrous", ctx.h, C.byref(prm), im[0].h, im[1].h, im[2].h, im[3].h, im[4].h, C.byref(C.c_int()))


@pytest.mark.parametrize("mode", [L.ATROUS_CONFIG, L.ATROUS_REFERENCE])
def test_atrous_full_1080p_matches_oracle(ctx, mode):
    """Both modes at the configured 1920x1080, 5 iterations (strides 1-16: LDS sub-lattice
    windows up to stride 8, dense tiles at 16, XCD runs with a partial last round, edge tiles
    of 1080 = 67.5 x 16 rows): exact weights bit-equal to the oracle on every pixel; fast
    weights within 1e-3 per pixel (relative above 1)."""
    H, W = 1080, 1920
    rad, pos, nrm = _ref_inputs(H, W, 5)
    if mode == L.ATROUS_CONFIG:
        a = rad.astype(np.float16).view(np.uint16).copy()
        a[..., 3] = np.float16(0.05).view(np.uint16)
        gb = np.zeros((H, W, 4), np.float32)
        gb[..., :3] = nrm[..., :3]
        gb[..., 3] = np.where(pos[..., 3] < 0, 0.0, 2.0 + pos[..., 2])
        b = gb.astype(np.float16).view(np.uint16).copy()
        imgs = [core.Image(ctx, W, H, L.FMT_RGBA16F) for _ in range(4)]
        imgs[0].upload(a)
        imgs[1].upload(b)
        o = a
        for k in range(5):
            o = O.atrous_cfg(o, b, 1 << k)
        of = o.view(np.float16).astype(np.float32)
        for flags in (0, L.ATROUS_FAST_MATH):
            slot = C.c_int()
            prm = L.AtrousParams(5, mode, 1.0, 128.0, 4.0, flags)
            L.call("rrh_atrous", ctx.h, C.byref(prm), imgs[0].h, imgs[1].h, None, imgs[2].h, imgs[3].h, C.byref(slot))
            g = imgs[2 + slot.value].download()
            gf = g.view(np.float16).astype(np.float32)
            step = np.spacing(np.abs(o.view(np.float16))).astype(np.float32)
            assert (np.abs(gf - of) <= 1e-3 * np.maximum(1.0, np.abs(of)) + step).all()
            if flags == 0:
                assert (g == o).all(axis=2).mean() == 1.0
    else:
        imgs = [core.Image(ctx, W, H) for _ in range(5)]
        for im, v in zip(imgs[:3], (rad, pos, nrm)):
            im.upload(v)
        o = rad
        for k in range(5):
            o = O.atrous_ref(o, pos, nrm, 1 << k)
        for flags in (0, L.ATROUS_FAST_MATH):
            slot = C.c_int()
            prm = L.AtrousParams(5, mode, 0.25, 0.2, 1.0, flags)
            L.call("rrh_atrous", ctx.h, C.byref(prm), imgs[0].h, imgs[1].h, imgs[2].h, imgs[3].h, imgs[4].h,
                   C.byref(slot))
            g = imgs[3 + slot.value].download()
            assert (np.abs(g - o) <= 1e-3 * np.maximum(1.0, np.abs(o))).all()
            if flags == 0:
                assert _bits_equal_frac(g.reshape(-1, 4), o.reshape(-1, 4)) == 1.0

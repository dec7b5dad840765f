"""Parity at the configured sizes: the 1,164,288-triangle heightfield (seed 2) at 1920x1080.

* Path tracer, 1 spp, max_bounces 5 (the bench's frame), frames 0 and 1: every pixel of a
  1-in-97 sample bit for bit against the oracle.
* ReSTIR DI with 32 candidates, temporal reuse (M cap 20 x 32) and 5 neighbours in 30 px, two
  frames: every stage on every pixel bit for bit (each stage on the GPU's own inputs).
* A-trous config mode on that render, as the reference wires DI's output and G-buffer into the
  denoiser (Application.cpp:534-535): rrh_atrous_prepare bit for bit, then 5 iterations of the
  exact filter bit for bit and of the fast filter within 1e-3 (relative above 1) plus one fp16 step.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from research_amd import _lib as L
from research_amd import core
from tests.helpers import ocam

pytestmark = pytest.mark.gpu

W, H = 1920, 1080


@pytest.fixture(scope="module")
def heightfield(ctx):
    sa = core.generate_scene(L.GEN_HEIGHTFIELD, 2, 1.0)
    assert sa.tri_count == 1_164_288
    gs = core.Scene(ctx, sa)
    os_ = O.Scene.from_arrays(sa)
    cam = core.camera_look_at(sa.eye, sa.center, sa.fov, W, H)
    gs.set_camera(cam)
    return sa, gs, os_, cam


def _bits(a):
    return np.ascontiguousarray(a).reshape(len(a), -1).view(np.uint32)


def _assert_same(g, o, what):
    diff = np.flatnonzero(~np.all(_bits(g) == _bits(o), axis=1))
    assert diff.size == 0, (what, diff.size, diff[:8])


def test_path_tracer_full_1080p_pixel_sample(ctx, heightfield):
    sa, gs, os_, cam = heightfield
    img = core.Image(ctx, W, H)
    pix = np.arange(0, W * H, 97, dtype=np.uint32)
    for frame in (0, 1):
        p = L.PTParams(5, frame, 3, 0, 0.0, 0)
        L.call("rrh_pt_render", ctx.h, gs.h, C.byref(p), img.h)
        g = img.download().reshape(-1, 4)[pix]
        o, _ = os_.pt_render(ocam(cam), 5, frame, pix)
        _assert_same(g, o, f"pt frame {frame}")
        assert (g[:, 3] == 1).all() and np.isfinite(g).all()


def test_restir_di_and_atrous_full_1080p(ctx, heightfield):
    sa, gs, os_, cam = heightfield
    oc = ocam(cam)
    im = lambda fmt=L.FMT_RGBA32F: core.Image(ctx, W, H, fmt)
    gp, gn, out = im(), im(), im()
    res, uv = [im(), im()], [im(L.FMT_R32U), im(L.FMT_R32U)]
    M = 32
    for frame in range(2):
        cur, prev = frame & 1, 1 - (frame & 1)
        prm, oprm = L.DIParams(frame, M, 1, 1, 5, 30), O.DIParams(frame, M, 1, 1, 5, 30)
        L.call("rrh_di_initial", ctx.h, gs.h, C.byref(prm), gp.h, gn.h, res[cur].h, uv[cur].h)
        ogp, ogn, ors, ouv = os_.di_initial(oc, oprm)
        g_gp, g_gn = gp.download().reshape(-1, 4), gn.download().reshape(-1, 4)
        _assert_same(g_gp, ogp, "gbuffer position")
        _assert_same(g_gn, ogn, "gbuffer normal")
        _assert_same(res[cur].download().reshape(-1, 4), ors, "initial reservoir")
        assert np.array_equal(uv[cur].download().reshape(-1), ouv)
        if frame > 0:
            rc0, rcuv0 = res[cur].download().reshape(-1, 4), uv[cur].download().reshape(-1)
            rp, rpuv = res[prev].download().reshape(-1, 4), uv[prev].download().reshape(-1)
            L.call("rrh_di_temporal", ctx.h, gs.h, C.byref(prm), gp.h, gn.h, res[cur].h, uv[cur].h, res[prev].h,
                   uv[prev].h)
            orc, orcuv = os_.di_temporal(oc, oprm, g_gp, g_gn, rc0, rcuv0, rp, rpuv)
            _assert_same(res[cur].download().reshape(-1, 4), orc, "temporal reservoir")
            assert np.array_equal(uv[cur].download().reshape(-1), orcuv)
        L.call("rrh_di_spatial_shade", ctx.h, gs.h, C.byref(prm), gp.h, gn.h, res[cur].h, uv[cur].h, out.h)
        g_out = out.download().reshape(-1, 4)
        o_out = os_.di_spatial_shade(oc, oprm, g_gp, g_gn, res[cur].download().reshape(-1, 4),
                                     uv[cur].download().reshape(-1))
        _assert_same(g_out, o_out, "shaded output")
        assert np.isfinite(g_out).all()

    # config-mode A-trous inputs from this frame: fp16 radiance + luminance variance, fp16 normal + linear depth
    r16, g16 = im(L.FMT_RGBA16F), im(L.FMT_RGBA16F)
    L.call("rrh_atrous_prepare", ctx.h, gs.h, out.h, gp.h, gn.h, r16.h, g16.h)
    o_r16, o_g16 = O.atrous_prepare(g_out.reshape(H, W, 4), g_gp.reshape(H, W, 4), g_gn.reshape(H, W, 4),
                                    cam.position, cam.direction)
    a, b = r16.download().reshape(H, W, 4), g16.download().reshape(H, W, 4)
    assert np.array_equal(a, o_r16) and np.array_equal(b, o_g16)
    assert (a.view(np.float16)[..., 3] > 0).mean() > 0.3      # a noisy render: variance is present
    o = o_r16
    for k in range(5):
        o = O.atrous_cfg(o, o_g16, 1 << k, 1.0, 128.0, 4.0)
    of16 = o.view(np.float16)
    t0, t1 = im(L.FMT_RGBA16F), im(L.FMT_RGBA16F)
    for flags in (0, L.ATROUS_FAST_MATH):
        slot = C.c_int()
        prm = L.AtrousParams(5, L.ATROUS_CONFIG, 1.0, 128.0, 4.0, flags)
        L.call("rrh_atrous", ctx.h, C.byref(prm), r16.h, g16.h, None, t0.h, t1.h, C.byref(slot))
        g = (t0, t1)[slot.value].download().reshape(H, W, 4)
        if flags == 0:
            assert np.array_equal(g, o), (g != o).any(axis=2).sum()
        else:
            gf, of = g.view(np.float16).astype(np.float32), of16.astype(np.float32)
            step = np.spacing(np.abs(of16)).astype(np.float32)
            assert np.isfinite(gf).all()
            assert (np.abs(gf - of) <= 1e-3 * np.maximum(1.0, np.abs(of)) + step).all()

"""ReSTIR GI (restir_gi.slang cs_initial / cs_temporal / cs_spatial_shade, host order
ReSTIRGIPass.cpp:428-651) on the GPU against the oracle's restatement, stage by stage on the
same inputs (reservoir planes, history planes, shaded output) over frames of a moving camera."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from research_amd import _lib as L
from research_amd import core
from tests.helpers import ocam, restir_cull
from tests.test_gpu_restir_pt import CASES, _bits, _cams, _res_same_frac, _same_frac

pytestmark = pytest.mark.gpu


def _params(frame, M=6, temporal=1, spatial=1, direct=1):
    args = (frame, M, temporal, spatial, 4, 12, direct, 1.0, 1e-3)
    return L.RGIParams(*args), O.RGIParams(*args)


@pytest.mark.parametrize("kind,seed,scale,step,W,H", CASES)
@pytest.mark.parametrize("cull", [0, 1])
def test_restir_gi_stages_match_oracle(ctx, kind, seed, scale, step, W, H, cull):
    with restir_cull(ctx, cull):
        _restir_gi_stages_match_oracle(ctx, kind, seed, scale, step, W, H)


def _restir_gi_stages_match_oracle(ctx, kind, seed, scale, step, W, H):
    sa = core.generate_scene(kind, seed, scale)
    gs = core.Scene(ctx, sa)
    os_ = O.Scene.from_arrays(sa)
    n = W * H
    im = lambda: core.Image(ctx, W, H)
    gp, gn, direct, out = im(), im(), im(), im()
    res = [[im() for _ in range(3)], [im() for _ in range(3)]]
    hist = [[im(), im()], [im(), im()]]
    dres, duv = im(), core.Image(ctx, W, H, L.FMT_R32U)
    for k in range(2):
        L.call("rrh_rgi_clear", ctx.h, core.planes(res[k]), hist[k][0].h, hist[k][1].h)
    assert (hist[1][0].download()[..., 3] == -1).all() and not res[1][2].download().any()
    dl = lambda planes: np.stack([p.download().reshape(-1, 4) for p in planes])
    reused = 0
    for f, cam in enumerate(_cams(sa, W, H, 4, step)):
        cur, prev = f & 1, 1 - (f & 1)
        gs.set_camera(cam)
        oc = ocam(cam)
        prm, oprm = _params(f)
        dprm = L.DIParams(f, 8, 1, 1, 4, 15)
        L.call("rrh_di_initial", ctx.h, gs.h, C.byref(dprm), gp.h, gn.h, dres.h, duv.h)
        L.call("rrh_di_spatial_shade", ctx.h, gs.h, C.byref(dprm), gp.h, gn.h, dres.h, duv.h, direct.h)
        g_gp, g_gn = gp.download().reshape(-1, 4), gn.download().reshape(-1, 4)
        L.call("rrh_rgi_initial", ctx.h, gs.h, C.byref(prm), gp.h, gn.h, core.planes(res[cur]), hist[cur][0].h,
               hist[cur][1].h)
        o_res, o_hp, o_hn = os_.rgi_initial(oc, oprm, g_gp, g_gn)
        g_res = dl(res[cur])
        assert _res_same_frac(g_res, o_res) == 1.0, f
        assert np.array_equal(_bits(hist[cur][0].download().reshape(-1, 4)), _bits(o_hp))
        assert np.array_equal(_bits(hist[cur][1].download().reshape(-1, 4)), _bits(o_hn))
        # the open heightfield: most one-bounce candidates escape to the sky (dropped, not counted)
        assert (o_res[2][:, 3].view(np.uint32) > 0).mean() > (0.01 if kind == L.GEN_HEIGHTFIELD else 0.3)
        if f > 0:
            g_prev = dl(res[prev])
            g_hp, g_hn = hist[prev][0].download().reshape(-1, 4), hist[prev][1].download().reshape(-1, 4)
            L.call("rrh_rgi_temporal", ctx.h, gs.h, C.byref(prm), gp.h, gn.h, core.planes(res[cur]),
                   core.planes(res[prev]), hist[prev][0].h, hist[prev][1].h)
            o_res = os_.rgi_temporal(oc, oprm, g_gp, g_gn, g_res, g_prev, g_hp, g_hn)
            g_res2 = dl(res[cur])
            assert _res_same_frac(g_res2, o_res) == 1.0, f
            reused += int((g_res2[2][:, 3].view(np.uint32) > g_res[2][:, 3].view(np.uint32)).sum())
            g_res = g_res2
        g_direct = direct.download().reshape(-1, 4)
        L.call("rrh_rgi_spatial_shade", ctx.h, gs.h, C.byref(prm), gp.h, gn.h, direct.h, core.planes(res[cur]),
               out.h)
        o_out = os_.rgi_spatial_shade(oc, oprm, g_gp, g_gn, g_res, direct=g_direct)
        g_out = out.download().reshape(-1, 4)
        assert _same_frac(g_out, o_out) == 1.0, f
        assert np.isfinite(g_out).all() and (g_out[:, :3] >= 0).all()
    assert reused > (0 if kind == L.GEN_HEIGHTFIELD else n // 10)


def test_restir_gi_stages_off_and_checks(ctx):
    sa = core.generate_scene(L.GEN_CORNELL_RANDOM, 2, 1.0)
    gs = core.Scene(ctx, sa)
    os_ = O.Scene.from_arrays(sa)
    W, H = 40, 32
    im = lambda fmt=L.FMT_RGBA32F: core.Image(ctx, W, H, fmt)
    gp, gn, out, dres, hp, hn = im(), im(), im(), im(), im(), im()
    duv = im(L.FMT_R32U)
    res = [im() for _ in range(3)]
    cam = core.camera_look_at((0.0, 0.3, 5.0), (0.0, 0.0, 0.0), sa.fov, W, H)
    gs.set_camera(cam)
    L.call("rrh_di_initial", ctx.h, gs.h, C.byref(L.DIParams(0, 4, 1, 1, 4, 15)), gp.h, gn.h, dres.h, duv.h)
    g_gp, g_gn = gp.download().reshape(-1, 4), gn.download().reshape(-1, 4)
    assert (g_gp[:, 3] < 0).any()
    prm, oprm = _params(5, M=8, temporal=0, spatial=0, direct=0)
    L.call("rrh_rgi_initial", ctx.h, gs.h, C.byref(prm), gp.h, gn.h, core.planes(res), hp.h, hn.h)
    L.call("rrh_rgi_spatial_shade", ctx.h, gs.h, C.byref(prm), gp.h, gn.h, None, core.planes(res), out.h)
    oc = ocam(cam)
    o_res, _, _ = os_.rgi_initial(oc, oprm, g_gp, g_gn)
    o_out = os_.rgi_spatial_shade(oc, oprm, g_gp, g_gn, o_res)
    assert _res_same_frac(np.stack([p.download().reshape(-1, 4) for p in res]), o_res) == 1.0
    g_out = out.download().reshape(-1, 4)
    assert _same_frac(g_out, o_out) == 1.0
    sky = g_gp[:, 3] < 0
    assert (g_out[sky, :3] > 0.09).all() and g_out[~sky, :3].max() > 0
    rc = L.RayCounts()
    L.call("rrh_rgi_ray_counts", ctx.h, C.byref(rc))
    assert rc.extension > W * H
    with pytest.raises(RuntimeError, match="three RGBA32F"):
        L.call("rrh_rgi_initial", ctx.h, gs.h, C.byref(prm), gp.h, gn.h, core.planes(res[:2] + [duv]), hp.h, hn.h)


@pytest.mark.parametrize("W,H", [(37, 29), (8, 1), (1, 9)])
def test_restir_gi_odd_extents_and_no_lights(ctx, W, H):
    from tests.test_oracle_restir_pt import furnace_box
    sa = furnace_box()
    gs = core.Scene(ctx, sa)
    os_ = O.Scene.from_arrays(sa)
    im = lambda: core.Image(ctx, W, H)
    gp, gn, out = im(), im(), im()
    res = [[im() for _ in range(3)], [im() for _ in range(3)]]
    hist = [[im(), im()], [im(), im()]]
    for k in range(2):
        L.call("rrh_rgi_clear", ctx.h, core.planes(res[k]), hist[k][0].h, hist[k][1].h)
    dl = lambda planes: np.stack([p.download().reshape(-1, 4) for p in planes])
    for f, cam in enumerate(_cams(sa, W, H, 2, 0.02)):
        cur, prev = f & 1, 1 - (f & 1)
        gs.set_camera(cam)
        oc = ocam(cam)
        g_gp, g_gn, _ = os_.gbuffer(oc)
        gp.upload(g_gp)
        gn.upload(g_gn)
        prm, oprm = _params(f, M=3, direct=0)
        L.call("rrh_rgi_initial", ctx.h, gs.h, C.byref(prm), gp.h, gn.h, core.planes(res[cur]), hist[cur][0].h,
               hist[cur][1].h)
        o_res, _, _ = os_.rgi_initial(oc, oprm, g_gp, g_gn)
        g_res = dl(res[cur])
        assert _res_same_frac(g_res, o_res) == 1.0
        if f > 0:
            g_prev = dl(res[prev])
            g_hp, g_hn = hist[prev][0].download().reshape(-1, 4), hist[prev][1].download().reshape(-1, 4)
            L.call("rrh_rgi_temporal", ctx.h, gs.h, C.byref(prm), gp.h, gn.h, core.planes(res[cur]),
                   core.planes(res[prev]), hist[prev][0].h, hist[prev][1].h)
            o_res = os_.rgi_temporal(oc, oprm, g_gp, g_gn, g_res, g_prev, g_hp, g_hn)
            g_res = dl(res[cur])
            assert _res_same_frac(g_res, o_res) == 1.0
        L.call("rrh_rgi_spatial_shade", ctx.h, gs.h, C.byref(prm), gp.h, gn.h, None, core.planes(res[cur]), out.h)
        o_out = os_.rgi_spatial_shade(oc, oprm, g_gp, g_gn, g_res)
        assert _same_frac(out.download().reshape(-1, 4), o_out) == 1.0

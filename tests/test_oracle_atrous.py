"""Independent pins of the oracle's A-trous edge-stopping weights.

Single-neighbour images: every pixel but the centre c and one tap q is a miss (reference mode)
or has its normal turned away (config mode: the normal weight is then exactly zero), so the
filtered centre is the closed-form weighted mean of c and q.  The weights are written here in
float64 from their definitions:
  * reference mode (atrous.slang:58-105, the edge-stopping function of Dammertz et al. 2010
    as the shader adapts it): B3 taps h(x) h(y) / 256 with h = {1, 4, 6, 4, 1}, clamp-to-edge
    taps, misses skipped, w = h exp(-|dp|^2 / sp^2) exp(-max(0, 1 - nc.nq) / sn^2)
    exp(-dl^2 / sl^2), Rec.601 luminance, sigma^2 floored at 1e-6, alpha from the centre;
  * config mode (Schied et al. 2017, eqs. 3-5): w_z = exp(-|zc - zq| / (sz |grad z . (c - q)| + 1e-3)),
    w_n = max(0, nc.nq)^sn, w_l = exp(-|lc - lq| / (sl sqrt(g3x3(var)) + 1e-10)), Rec.709
    luminance, grad z by central differences, variance out = sum w^2 var / (sum w)^2, fp16 I/O.
"""
import numpy as np
import pytest

from oracle import oracle as O

KW = {-2: 1.0, -1: 4.0, 0: 6.0, 1: 4.0, 2: 1.0}


def luma601(c):
    return 0.299 * c[0] + 0.587 * c[1] + 0.114 * c[2]


def luma709(c):
    return 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2]


def ref_weight(pc, nc, lc, pq, nq, lq, sp, sn, sl):
    nc = nc / np.linalg.norm(nc)
    nq = nq / np.linalg.norm(nq)
    return (np.exp(-np.sum((pq - pc) ** 2) / max(sp * sp, 1e-6)) *
            np.exp(-max(0.0, 1.0 - nc @ nq) / max(sn * sn, 1e-6)) *
            np.exp(-(lq - lc) ** 2 / max(sl * sl, 1e-6)))


def ref_images(W, H):
    rad = np.zeros((H, W, 4), np.float32)
    pos = np.zeros((H, W, 4), np.float32)
    pos[..., 3] = -1.0
    nrm = np.zeros((H, W, 4), np.float32)
    return rad, pos, nrm


@pytest.mark.parametrize("stride,cx,cy,tx,ty", [
    (1, 20, 15, 1, 0), (2, 20, 15, -2, 1), (4, 20, 15, 2, 2), (8, 20, 40, 0, -2), (16, 40, 40, 2, -1),
    (4, 0, 40, 1, 0),      # centre on the left edge: taps -2, -1, 0 all clamp onto it
    (2, 0, 0, 1, 1),       # corner: the centre collects 11 x 11 / 256
    (8, 79, 79, -1, -2),   # right / bottom corner
])
@pytest.mark.parametrize("sig", [(0.25, 0.2, 1.0), (1.0, 0.5, 0.3)])
def test_reference_mode_single_neighbour_closed_form(stride, cx, cy, tx, ty, sig):
    W = H = 80
    sp, sn, sl = sig
    rad, pos, nrm = ref_images(W, H)
    qx, qy = cx + tx * stride, cy + ty * stride
    c_rgb, q_rgb = np.array([1.5, 0.8, 0.3]), np.array([0.2, 1.1, 2.4])
    rad[cy, cx] = [*c_rgb, 0.75]
    rad[qy, qx] = [*q_rgb, 0.1]
    pos[cy, cx] = [0.3, 0.1, -0.2, 1.0]
    pos[qy, qx] = [0.35, 0.12, -0.15, 2.0]
    nrm[cy, cx] = [0.0, 2.0, 0.0, 0.0]          # not unit: the filter normalises
    nrm[qy, qx] = [0.3, 1.0, 0.2, 0.0]
    out = O.atrous_ref(rad, pos, nrm, stride, sp, sn, sl)
    # the taps that land on c and on q after clamping
    wc = wq = 0.0
    for ty_ in range(-2, 3):
        for tx_ in range(-2, 3):
            x = min(max(cx + tx_ * stride, 0), W - 1)
            y = min(max(cy + ty_ * stride, 0), H - 1)
            h = KW[tx_] * KW[ty_] / 256.0
            if (x, y) == (cx, cy):
                wc += h
            elif (x, y) == (qx, qy):
                wq += h
    rc, rq = rad[cy, cx].astype(np.float64), rad[qy, qx].astype(np.float64)
    e = ref_weight(pos[cy, cx, :3].astype(np.float64), nrm[cy, cx, :3].astype(np.float64), luma601(rc),
                   pos[qy, qx, :3].astype(np.float64), nrm[qy, qx, :3].astype(np.float64), luma601(rq), sp, sn, sl)
    want = (wc * rc[:3] + wq * e * rq[:3]) / (wc + wq * e)
    assert wq > 0 and 1e-3 < wq * e / (wc + wq * e) < 0.999     # both terms matter (well above rtol)
    assert np.allclose(out[cy, cx, :3], want, rtol=1e-5), (out[cy, cx], want)
    assert out[cy, cx, 3] == np.float32(0.75)
    # a miss centre is copied unchanged
    assert np.array_equal(out[0, 5], rad[0, 5]) or (cx, cy) == (5, 0)


def f16(x):
    return np.asarray(x, np.float16)


def cfg_images(W, H, z0, gx, gy, var):
    yy, xx = np.mgrid[0:H, 0:W]
    gbuf = np.zeros((H, W, 4), np.float16)
    gbuf[..., 2] = -1.0                      # turned away from the centre's +z: w_n = 0
    gbuf[..., 3] = z0 + gx * xx + gy * yy    # exactly representable plane
    assert np.array_equal(gbuf[..., 3].astype(np.float64), z0 + gx * xx + gy * yy)
    rad = np.zeros((H, W, 4), np.float16)
    rad[..., :3] = f16([0.5, 0.5, 0.5])
    rad[..., 3] = var
    return rad, gbuf


@pytest.mark.parametrize("stride,tx,ty", [(1, 1, 0), (2, -1, 1), (4, 1, 1), (8, 0, -1), (8, -2, 1)])
@pytest.mark.parametrize("nq,var", [((0.0, 0.1, 1.0), 0.0625), ((0.05, -0.04, 1.0), 0.25)])
@pytest.mark.parametrize("sn", [128.0, 100.0])   # 2^7 by squaring; otherwise exp(sn log x)
def test_config_mode_single_neighbour_closed_form(stride, tx, ty, nq, var, sn):
    W, H, cx, cy = 48, 40, 20, 18
    z0, gx, gy = 4.0, 1.0 / 16, 1.0 / 32
    sz, sl = 1.0, 4.0
    rad, gbuf = cfg_images(W, H, z0, gx, gy, var)
    qx, qy = cx + tx * stride, cy + ty * stride
    gbuf[cy, cx, :3] = f16([0.0, 0.0, 1.0])
    n_q = np.array(nq) / np.linalg.norm(nq)
    gbuf[qy, qx, :3] = f16(n_q)
    rad[cy, cx, :3] = f16([1.0, 0.5, 0.25])
    rad[qy, qx, :3] = f16([0.25, 0.75, 1.5])
    out = O.atrous_cfg(rad.view(np.uint16), gbuf.view(np.uint16), stride, sz, sn, sl).view(np.float16)
    c = rad[cy, cx].astype(np.float64)
    q = rad[qy, qx].astype(np.float64)
    nc, nqd = gbuf[cy, cx, :3].astype(np.float64), gbuf[qy, qx, :3].astype(np.float64)
    zc, zq = float(gbuf[cy, cx, 3]), float(gbuf[qy, qx, 3])
    dz = gx * tx * stride + gy * ty * stride           # grad z . offset (central differences are exact here)
    w_z = abs(zc - zq) / (sz * abs(dz) + 1e-3)
    w_l = abs(luma709(c) - luma709(q)) / (sl * np.sqrt(var) + 1e-10)
    w_n = max(0.0, nc @ nqd) ** sn
    wc = 36.0 / 256.0
    wq = KW[tx] * KW[ty] / 256.0 * np.exp(-w_l - w_z) * w_n
    assert 5e-3 < wq / (wc + wq) < 0.999          # far above one fp16 rounding
    want_rgb = (wc * c[:3] + wq * q[:3]) / (wc + wq)
    want_var = (wc * wc * var + wq * wq * var) / (wc + wq) ** 2
    got = out[cy, cx].astype(np.float64)
    tol = lambda v: np.abs(v) * 2.0 ** -10 + 2.0 ** -24       # one fp16 rounding of the result
    assert (np.abs(got[:3] - want_rgb) <= 1.5 * tol(want_rgb)).all(), (got, want_rgb)
    assert abs(got[3] - want_var) <= 1.5 * tol(want_var), (got[3], want_var)


def test_config_mode_miss_tap_and_miss_centre():
    W, H, cx, cy = 32, 32, 10, 10
    rad, gbuf = cfg_images(W, H, 4.0, 1.0 / 16, 1.0 / 32, 0.0625)
    gbuf[cy, cx, :3] = f16([0.0, 0.0, 1.0])
    gbuf[cy, cx + 2, :3] = f16([0.0, 0.0, 1.0])
    gbuf[cy, cx + 2, 3] = 0.0                   # the only facing tap is a miss: skipped
    rad[cy, cx, :3] = f16([1.0, 0.5, 0.25])
    gbuf[3, 3, 3] = -1.0                        # a miss centre is copied
    out = O.atrous_cfg(rad.view(np.uint16), gbuf.view(np.uint16), 2).view(np.float16)
    assert np.array_equal(out[cy, cx, :3], rad[cy, cx, :3])
    assert out[cy, cx, 3] == rad[cy, cx, 3]     # w_c^2 var / w_c^2
    assert np.array_equal(out[3, 3], rad[3, 3])


def test_prepare_variance_depth_and_packing():
    """rrh_atrous_prepare's restatement: alpha = population variance of Rec.709 luminance over the
    valid pixels of the clamped 3x3 window (edge pixels repeat at the borders), depth = (p - eye) .
    forward (0 on a miss), everything rounded once to fp16."""
    rng = np.random.default_rng(4)
    H, W = 23, 31
    rad = rng.uniform(0.0, 2.0, (H, W, 4)).astype(np.float32)
    gpos = np.zeros((H, W, 4), np.float32)
    gpos[..., :3] = rng.uniform(-3, 3, (H, W, 3))
    gpos[..., 3] = 1.0
    gpos[rng.random((H, W)) < 0.2, 3] = -1.0
    gpos[0:3, 0:3, 3] = -1.0                   # a corner whose window is all misses
    gnrm = rng.normal(size=(H, W, 4)).astype(np.float32)
    eye = np.array([0.5, -0.2, 6.0])
    fwd = np.array([0.1, 0.2, -1.0]) / np.linalg.norm([0.1, 0.2, -1.0])
    r16, g16 = O.atrous_prepare(rad, gpos, gnrm, eye, fwd)
    r16, g16 = r16.view(np.float16), g16.view(np.float16)
    lum = rad[..., :3].astype(np.float64) @ np.array([0.2126, 0.7152, 0.0722])
    want_var = np.zeros((H, W))
    for y in range(H):
        for x in range(W):
            ls = [lum[min(max(y + dy, 0), H - 1), min(max(x + dx, 0), W - 1)]
                  for dy in (-1, 0, 1) for dx in (-1, 0, 1)
                  if gpos[min(max(y + dy, 0), H - 1), min(max(x + dx, 0), W - 1), 3] >= 0]
            want_var[y, x] = np.var(ls) if ls else 0.0
    got_var = r16[..., 3].astype(np.float64)
    assert np.all(np.abs(got_var - want_var) <= want_var * 2.0 ** -10 + 2e-6)
    assert (got_var[0:2, 0:2] == 0).all()
    assert np.array_equal(r16[..., :3], rad[..., :3].astype(np.float16))     # one RNE rounding
    assert np.array_equal(g16[..., :3], gnrm[..., :3].astype(np.float16))
    z = np.where(gpos[..., 3] >= 0, (gpos[..., :3].astype(np.float64) - eye) @ fwd, 0.0)
    got_z = g16[..., 3].astype(np.float64)
    assert np.all(np.abs(got_z - z) <= np.abs(z) * 2.0 ** -10 + 1e-6)
    assert (got_z[gpos[..., 3] < 0] == 0).all()

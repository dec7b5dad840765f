"""Independent pins of the oracle's ReSTIR DI merges (restir_di.slang:186-472).

Closed forms of the reservoir arithmetic, not a restatement of it:
  * temporal reuse of a reservoir with itself (same light, same surface) doubles M and w_sum
    and keeps W, because W = w_sum / (M p_hat) (:254-258, :290-334);
  * the M cap (:296-316): with self-reuse, M -> min(M + min(M_prev, cap), cap) and
    w_sum -> w_sum M_new / M, W unchanged; the recomputed W is capped at 1e4 (:330);
  * with a single light every reservoir holds it and every merge keeps W = 1, so the shaded
    output is exactly f E cos V of that light (:386-471), checked against a float64 BSDF
    (tests/bsdf64.py) and the oracle's occlusion test;
  * with temporal and spatial reuse off, resampled importance sampling is unbiased: the mean
    over frames equals the brute-force visible direct lighting summed over all lights
    (point and directional branches of compute_p_hat, :152-178, and of the shading, :432-448).
"""
import numpy as np
import pytest

from oracle import oracle as O
from research_amd import _lib as L
from research_amd import core
from tests.bsdf64 import cook_torrance, normalize
from tests.helpers import ocam

EMPTY = 0xFFFFFFFF


def u32(a):
    return np.ascontiguousarray(a).view(np.uint32)


def f32_of(u):
    return np.asarray(u, np.uint32).view(np.float32)


def cornell(W=48, H=36, lights=None, mats=None):
    sa = core.generate_scene(L.GEN_CORNELL_REF)
    if lights is not None:
        sa.lights = lights
    if mats is not None:
        sa.mats = mats
    cam = core.camera_look_at(sa.eye, sa.center, sa.fov, W, H)
    return sa, O.Scene.from_arrays(sa), ocam(cam)


def light(pos=(0, 0, 0), typ=0, direction=(0, 0, 0), emission=(1, 1, 1), intensity=1.0):
    l = np.zeros(1, L.LIGHT)
    l["position"], l["type"], l["direction"] = pos, typ, direction
    l["emission"], l["intensity"] = emission, intensity
    return l


def valid_reservoirs(gp, rs):
    return (gp[:, 3] >= 0) & (u32(rs[:, 0]) != EMPTY) & (u32(rs[:, 3]) > 0)


def test_temporal_self_reuse_doubles_m_and_wsum_keeps_w():
    sa, os_, oc = cornell()
    prm = O.DIParams(0, 16, 1, 1, 5, 30)
    gp, gn, rs, uv = os_.di_initial(oc, prm)
    v = valid_reservoirs(gp, rs)
    assert v.sum() > 500
    rc, rcuv = os_.di_temporal(oc, prm, gp, gn, rs, uv, rs, uv)
    assert np.array_equal(u32(rc[v, 0]), u32(rs[v, 0]))             # same light
    assert np.array_equal(u32(rc[v, 3]), 2 * u32(rs[v, 3]))         # M doubles (32 < cap 320)
    assert np.allclose(rc[v, 1], 2 * rs[v, 1], rtol=1e-6)            # w_sum doubles
    assert np.allclose(rc[v, 2], rs[v, 2], rtol=1e-6)                # W kept
    # pixels without history (miss or empty reservoir) are left bit for bit
    assert np.array_equal(u32(rc[~v]), u32(rs[~v]))


@pytest.mark.parametrize("m_prev", [1, 200, 310, 1000, 70000])
def test_temporal_m_cap_closed_form(m_prev):
    sa, os_, oc = cornell()
    M = 16
    cap = 20 * M
    prm = O.DIParams(3, M, 1, 1, 5, 30)
    gp, gn, rs, uv = os_.di_initial(oc, prm)
    v = valid_reservoirs(gp, rs)
    rp = rs.copy()
    rp[:, 3] = f32_of(np.full(len(rp), m_prev, np.uint32))
    rc, _ = os_.di_temporal(oc, prm, gp, gn, rs, uv, rp, uv)
    m_new = min(M + min(m_prev, cap), cap)
    assert (u32(rc[v, 3]) == m_new).all()
    assert np.allclose(rc[v, 1], rs[v, 1] * (m_new / M), rtol=2e-6)
    assert np.allclose(rc[v, 2], rs[v, 2], rtol=2e-6)


def test_temporal_w_is_capped_at_1e4():
    sa, os_, oc = cornell()
    prm = O.DIParams(0, 16, 1, 1, 5, 30)
    gp, gn, rs, uv = os_.di_initial(oc, prm)
    v = valid_reservoirs(gp, rs)
    rp = rs.copy()
    rp[:, 2] = 1e6      # an exploding history weight
    rc, _ = os_.di_temporal(oc, prm, gp, gn, rs, uv, rp, uv)
    # uncapped W would be (16 W + 1e6 M_prev) / (16 + M_prev) > 1e4
    assert (rc[v, 2] == np.float32(1e4)).all()


def brute_direct(sa, os_, oc, gp, gn, lights):
    """float64 sum over lights of f E cos V at each G-buffer surface (0 for misses/emitters)."""
    P = len(gp)
    hit = gp[:, 3] >= 0
    mat = u32(gn[:, 3]).copy()
    mat[~hit] = 0
    m = sa.mats[mat]
    emissive = m["emissive"].max(1) > 0
    surf = hit & ~emissive
    p, n = gp[:, :3].astype(np.float64), gn[:, :3].astype(np.float64)
    campos = np.array(oc.position, np.float64)
    wo = normalize(campos - p)
    total = np.zeros((P, 3))
    for lt in lights:
        if lt["type"] == 0:
            to = lt["position"].astype(np.float64) - p
            dist = np.linalg.norm(to, axis=1)
            l = to / np.maximum(dist, 1e-6)[:, None]
            irr = (lt["emission"] * lt["intensity"])[None, :] / np.maximum(dist * dist, 1e-6)[:, None]
        else:
            l = np.tile(normalize(-lt["direction"].astype(np.float64)), (P, 1))
            dist = np.full(P, 1e9)
            irr = np.tile(lt["emission"] * lt["intensity"], (P, 1)).astype(np.float64)
        ndl = np.maximum(np.sum(n * l, 1), 0.0)
        rays = np.zeros((P, 8), np.float32)
        rays[:, :3] = p + n * 1e-4
        rays[:, 3] = 0.001
        rays[:, 4:7] = l
        rays[:, 7] = dist - 0.01
        vis = u32(os_.trace(rays)[:, 1]) == EMPTY
        f = cook_torrance(n, wo, l, m["base_color"][:, :3], m["metallic"], m["roughness"])
        total += np.where((surf & (ndl > 0) & vis)[:, None], f * irr * ndl[:, None], 0.0)
    return total, surf


def test_single_light_every_merge_keeps_w_one():
    sa, os_, oc = cornell(64, 48)
    lights = sa.lights
    assert len(lights) == 1 and lights[0]["type"] == 0
    prev = None
    for frame in range(3):
        prm = O.DIParams(frame, 16, 1, 1, 5, 30)
        gp, gn, rs, uv = os_.di_initial(oc, prm)
        v = valid_reservoirs(gp, rs)
        assert np.allclose(rs[v, 2], 1.0, rtol=1e-6)                # initial: W = 1
        if prev is not None:
            pr = prev[0]
            had = (u32(pr[:, 0]) != EMPTY) & (u32(pr[:, 3]) > 0)
            m_want = np.where(had, np.minimum(16 + np.minimum(u32(pr[:, 3]), 320), 320), 16)
            rs, uv = os_.di_temporal(oc, prm, gp, gn, rs, uv, *prev)
            assert np.allclose(rs[v, 2], 1.0, rtol=1e-6)            # temporal (M capped at 320)
            assert (u32(rs[v, 3]) == m_want[v]).all()
        out = os_.di_spatial_shade(oc, prm, gp, gn, rs, uv)
        want, surf = brute_direct(sa, os_, oc, gp, gn, lights)
        # spatial merges (5 neighbours in 30 px) also keep W = 1: output = f E cos V
        assert np.allclose(out[surf, :3], want[surf], rtol=2e-5, atol=1e-6), np.abs(out[surf, :3] - want[surf]).max()
        prev = (rs, uv)


def test_initial_resampling_is_unbiased_over_frames():
    """Temporal and spatial off: E[f E cos V W] over frames = the sum over every light."""
    ls = np.concatenate([
        light((0.0, 0.9, 0.0), emission=(15, 15, 12)),
        light((-0.7, 0.3, 0.6), emission=(1, 4, 8), intensity=0.5),
        light((0.6, -0.6, 0.2), emission=(6, 2, 1)),
        light((0.2, 0.5, -0.8), emission=(0.2, 0.2, 0.2), intensity=20.0),
        light(typ=1, direction=(-0.2, -0.3, -1.0), emission=(3, 3, 2.5)),   # through the open front
    ])
    mats = core.generate_scene(L.GEN_CORNELL_REF).mats.copy()
    mats["metallic"] = [0.0, 0.6, 0.2, 0.0]
    mats["roughness"] = [0.9, 0.35, 0.5, 1.0]
    sa, os_, oc = cornell(24, 18, lights=ls, mats=mats)
    F = 300
    s1 = s2 = 0.0
    acc_d = np.zeros((24 * 18, 3))
    acc_d2 = np.zeros((24 * 18, 3))
    acc_bf = np.zeros((24 * 18, 3))
    for frame in range(F):
        prm = O.DIParams(frame, 4, 0, 0, 5, 30)
        gp, gn, rs, uv = os_.di_initial(oc, prm)
        out = os_.di_spatial_shade(oc, prm, gp, gn, rs, uv)
        bf, surf = brute_direct(sa, os_, oc, gp, gn, ls)
        d = np.where(surf[:, None], out[:, :3] - bf, 0.0)
        acc_d += d
        acc_d2 += d * d
        acc_bf += bf
    mean_d = acc_d / F
    se = np.sqrt(np.maximum(acc_d2 / F - mean_d ** 2, 0.0) / F)
    # global relative bias within 4 standard errors, and no pixel beyond 5.5
    tot, tot_se = acc_d.sum() / F, np.sqrt((se ** 2).sum())
    assert abs(tot) < 4 * tot_se, (tot, tot_se, acc_bf.sum() / F)
    z = np.abs(mean_d) / np.maximum(se, 1e-12)
    assert (z[se > 0] < 5.5).all(), z.max()
    assert acc_bf.sum() / F > 10 * tot_se   # the test has power: signal far above the noise

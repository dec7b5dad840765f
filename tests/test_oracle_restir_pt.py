"""Pins of the oracle's ReSTIR PT and ReSTIR GI restatements (restir_pt.slang, restir_gi.slang) by
what the mathematics fixes:

* furnace: inside a closed box whose Lambertian walls (albedo rho) all emit Le and no light is
  sampled, every path suffix is exactly Le (continued bounces add no emission), so the
  resampled estimate f cos Le W of every reservoir equals rho Le whatever candidate was kept
  (cosine source pdf) -- a wrong weight, M, pdf or W formula breaks it;
* with a guiding lobe the source pdf varies and candidates below the surface are dropped
  before they are counted, which conditions the estimate on acceptance: the image mean is the
  closed form rho Le (1 - (1 - P)^M) / P (a wrong MIS pdf moves it);
* temporal reuse of a reservoir with itself (static camera, history = current frame) finds
  every history pixel, doubles M (up to the cap) and keeps W wherever the visibility ray
  reaches the reconnection vertex: the reconnection Jacobian of a sample to its own primary
  surface is exactly one; where the visibility ray's end crosses the vertex's own surface
  (closed form from the ray offsets) the reuse is rejected."""
import numpy as np
import pytest

from oracle import oracle as O
from research_amd import _lib as L
from research_amd import core
from tests.helpers import ocam

RHO, LE = 0.5, 1.0


def furnace_box(rho=RHO, le=LE, outward=False):
    """Cube [-1, 1]^3 of 12 triangles facing inward (or wound outward: geometric normals out,
    vertex normals still in), one pure-Lambert emissive material."""
    tris = []
    for a in range(3):
        b, c = (a + 1) % 3, (a + 2) % 3
        for s in (-1.0, 1.0):
            q = []
            for u, v in ((-1, -1), (1, -1), (1, 1), (-1, 1)):
                p = [0.0, 0.0, 0.0]
                p[a], p[b], p[c] = s, u, v
                q.append(p)
            inward = np.zeros(3)
            inward[a] = -s
            for t in ((q[0], q[1], q[2]), (q[0], q[2], q[3])):
                t = np.array(t)
                if (np.dot(np.cross(t[1] - t[0], t[2] - t[0]), inward) < 0) != outward:
                    t = t[[0, 2, 1]]
                tris.append((t, inward))
    v = np.zeros(len(tris) * 3, L.VERTEX)
    for k, (t, n) in enumerate(tris):
        v["pos"][3 * k:3 * k + 3] = t
        v["normal"][3 * k:3 * k + 3] = n
    v["tangent"] = [1, 0, 0, 1]
    idx = np.arange(len(v), dtype=np.uint32)
    mesh = np.zeros(1, L.MESH)
    mesh["vertex_count"], mesh["index_count"] = len(v), len(idx)
    mat = np.zeros(1, L.MATERIAL)
    mat["base_color"], mat["roughness"], mat["ior"] = [rho, rho, rho, 1.0], 1.0, 1.5
    mat["emissive"] = le
    mat["flags"] = L.MAT_PURE_LAMBERT
    inst = np.zeros(1, L.INSTANCE)
    inst["transform"] = np.eye(4, dtype=np.float32).reshape(-1)
    return core.SceneArrays(v, idx, mesh, mat, inst, np.zeros(0, L.LIGHT), 0, (0.3, 0.2, 0.9), (0.0, 0.0, 0.0), 60.0)


@pytest.fixture()
def culled():
    """Rays cull back faces (RAY_FLAG_CULL_BACK_FACING_TRIANGLES without the instances'
    CULL_DISABLE): at ray bias 0 only culling keeps a ray from re-hitting the wall it leaves."""
    O.lib().orc_set_restir_cull(1)
    yield
    O.lib().orc_set_restir_cull(0)


def _setup(W=48, H=36):
    sa = furnace_box()
    os_ = O.Scene.from_arrays(sa)
    cam = core.camera_look_at(sa.eye, sa.center, sa.fov, W, H)
    oc = ocam(cam)
    gp, gn, _ = os_.gbuffer(oc)
    assert (gp[:, 3] == 1).all()  # every pixel sees a wall
    return os_, oc, gp, gn


def _prm(M, bounces, guiding=0, temporal=1, bias=0.0):
    """bias 0 (with culled rays): the walls face inward, so no ray can hit the surface it
    leaves, and the sampled direction is exactly the one toward the hit (with a bias the offset
    origin tilts it by about bias / distance, and rays leak past corners)."""
    return O.RPTParams(3, M, bounces, temporal, 0, 4, 12, 0, 1.0, bias, guiding, 0.5)


def _estimate(gp, gn, res):
    """f cos Le W of each reservoir with the Lambert BRDF rho / pi."""
    n = gn[:, :3] / np.linalg.norm(gn[:, :3], axis=1, keepdims=True)
    to = res[0][:, :3].astype(np.float64) - gp[:, :3]
    wi = to / np.linalg.norm(to, axis=1, keepdims=True)
    cos = np.maximum((n * wi).sum(1), 0)
    return RHO / np.pi * cos * res[2][:, 0] * res[2][:, 3]


@pytest.mark.parametrize("M,bounces", [(1, 1), (8, 1), (8, 3)])
def test_furnace_every_reservoir_estimates_rho_le(culled, M, bounces):
    os_, oc, gp, gn = _setup()
    res, hp, hn = os_.rpt_initial(oc, _prm(M, bounces), gp, gn)
    assert (res[3][:, 3] == 1).all()
    assert (res[1][:, 3].view(np.uint32) == M).all()          # every candidate counted
    assert (res[2][:, :3] == np.float32(LE)).all()            # suffix = Le exactly
    assert np.allclose(_estimate(gp, gn, res), RHO * LE, rtol=2e-5)
    assert np.array_equal(hp, gp) and np.array_equal(hn, gn)  # history = this frame's surface


def test_ray_bias_leaks_only_at_corners():
    os_, oc, gp, gn = _setup()
    res, _, _ = os_.rpt_initial(oc, _prm(1, 1, bias=1e-3), gp, gn)
    lost = res[3][:, 3] < 0.5
    assert 0 < lost.sum() < 0.01 * len(lost)
    # within a bias of a second wall: the offset origin crossed it
    near = (np.abs(np.abs(gp[lost, :3]) - 1) < 1e-3).sum(1) >= 2
    assert near.all()


@pytest.mark.parametrize("M,bounces", [(1, 1), (8, 3)])
def test_two_sided_rays_ignore_winding(M, bounces):
    """The reference traces two-sided (CULL_DISABLE instances, AccelStructure.cpp:90): a box
    wound outward gives the same reservoirs as one wound inward, up to the rounding of the
    swapped barycentrics, and the furnace estimate rho Le holds away from the corners (a bias
    of 1e-3 tilts a direction by about bias / distance).  Culled rays leave the outward-wound
    box through its walls: every candidate escapes and is dropped before it is counted."""
    prm = _prm(M, bounces, bias=1e-3)
    os_in, oc, gp, gn = _setup()
    os_out = O.Scene.from_arrays(furnace_box(outward=True))   # same surfaces, G-buffer of the inward box
    res_in, _, _ = os_in.rpt_initial(oc, prm, gp, gn)
    res_out, _, _ = os_out.rpt_initial(oc, prm, gp, gn)
    ok = (res_in[3][:, 3] == 1) & (res_out[3][:, 3] == 1)
    assert ok.mean() > 0.99
    assert (res_in[1][ok, 3].view(np.uint32) == res_out[1][ok, 3].view(np.uint32)).mean() > 0.99
    for k in range(3):
        assert np.allclose(res_in[k][ok], res_out[k][ok], rtol=1e-4, atol=1e-4)
    # the offset origin moves cos toward the hit by about bias / (distance cos): keep reconnections
    # at distance > .5 and cos > .2 (relative error < 1e-2)
    to = res_out[0][:, :3] - gp[:, :3]
    dist = np.linalg.norm(to, axis=1)
    cos = (gn[:, :3] * to).sum(1) / np.maximum(dist, 1e-9)
    far = ok & (dist > 0.5) & (cos > 0.2)
    assert far.mean() > 0.5
    d = np.abs(_estimate(gp, gn, res_out)[far] / (RHO * LE) - 1)
    # one candidate: every pixel within 1e-2; eight: W also carries the others' (grazing) tilt
    assert (d.max() < 1e-2) if M == 1 else (np.median(d) < 2e-3 and np.quantile(d, 0.9) < 6e-3), d
    O.lib().orc_set_restir_cull(1)
    try:
        res_cull, _, _ = os_out.rpt_initial(oc, prm, gp, gn)
    finally:
        O.lib().orc_set_restir_cull(0)
    assert (res_cull[3][:, 3] == 0).all() and (res_cull[1][:, 3].view(np.uint32) == 0).all()


def _fib_sphere(n):
    i = np.arange(n) + 0.5
    z = 1 - 2 * i / n
    r = np.sqrt(1 - z * z)
    phi = np.pi * (1 + 5 ** 0.5) * i
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)


def test_furnace_guided_sampling_matches_its_closed_form(culled):
    """One-sample MIS with a vMF lobe: candidates whose direction falls below the surface are
    dropped before they are counted in M (restir_pt.slang:618-619, 731-732), so the estimate
    is conditioned on acceptance: E = rho Le (1 - (1 - P)^M) / P per pixel, with P = alpha P_vmf(n . w
    > 0) + 1 - alpha.  P_vmf from the textbook vMF density integrated over a Fibonacci sphere.
    A wrong MIS pdf moves the mean off this value."""
    os_, oc, gp, gn = _setup(64, 48)
    count = 1 << 10
    mu, kappa, alpha, M = np.array([0.0, 0.6, 0.8]), 6.0, 0.5, 8
    fitted = np.zeros((count, 4), np.float32)
    fitted[:] = [*mu, kappa]  # one valid lobe in every cell
    res, _, _ = os_.rpt_initial(oc, _prm(M, 1, guiding=1), gp, gn, grid=(fitted, 0.1, 0.02, count))
    est = _estimate(gp, gn, res)
    assert not np.allclose(est[est > 0], RHO * LE, rtol=1e-3)  # the source pdf varies per candidate
    w = _fib_sphere(200000)
    dens = kappa / (2 * np.pi * (1 - np.exp(-2 * kappa))) * np.exp(kappa * (w @ mu - 1))
    n = gn[:, :3].astype(np.float64)
    p_up = {}
    for k in map(tuple, np.unique(n, axis=0)):
        p_up[k] = (dens * (w @ np.array(k) > 0)).sum() * 4 * np.pi / len(w)
    P = np.array([alpha * p_up[tuple(k)] + 1 - alpha for k in n])
    expect = RHO * LE * (1 - (1 - P) ** M) / P
    assert expect.mean() > 1.05 * RHO * LE  # the conditioning is visible at this lobe
    assert abs(est.mean() / expect.mean() - 1) < 0.015


def test_temporal_reuse_with_itself_keeps_w():
    os_, oc, gp, gn = _setup()
    prm = _prm(4, 2, bias=1e-3)  # the visibility ray stops a bias short of its target
    res, hp, hn = os_.rpt_initial(oc, prm, gp, gn)
    out = os_.rpt_temporal(oc, prm, gp, gn, res, res, hp, hn)
    M0, M1 = res[1][:, 3].view(np.uint32), out[1][:, 3].view(np.uint32)
    keep = (M1 == 2 * M0) & (M0 > 0)
    # trace_visibility (restir_pt.slang:449-458) leaves from p + n bias toward the target and
    # stops a bias short of it along the ray; its end lies bias (n . n_rc + cos_rc) in front
    # of the target's plane, so where the primary faces the reconnection vertex's surface more
    # than the ray meets it (opposite walls), the ray crosses that surface and the reuse is
    # rejected: the reference's behaviour, kept
    n = gn[:, :3].astype(np.float64)
    to = res[0][:, :3] - gp[:, :3].astype(np.float64)
    cos_rc = -(res[1][:, :3] * to).sum(1) / np.linalg.norm(to, axis=1)
    ahead = (n * res[1][:, :3]).sum(1) + cos_rc
    assert (ahead < -0.01).mean() > 0.05
    assert not keep[ahead < -0.01].any()
    assert keep[(ahead > 0.01) & (cos_rc > 0.06)].all()
    under = keep & (res[2][:, 3] < 50)  # W above 50 is capped (restir_pt.slang:838, 866)
    assert under.mean() > 0.7
    assert np.allclose(out[2][under, 3], res[2][under, 3], rtol=1e-5)
    assert np.allclose(out[0][under, 3], 2 * res[0][under, 3], rtol=1e-6)  # w_sum doubles
    assert (out[2][keep, 3] <= 50).all()
    # the cap: 20 x candidates
    big = res.copy()
    big[1][:, 3] = np.full(len(gp), 70, np.uint32).view(np.float32)
    out2 = os_.rpt_temporal(oc, prm, gp, gn, big, big, hp, hn)
    assert (out2[1][keep, 3].view(np.uint32) == 80).all()


# ---- ReSTIR GI (restir_gi.slang): the one-bounce sibling -------------------------------------
def _gi_prm(M, temporal=1, bias=1e-3):
    return O.RGIParams(3, M, temporal, 0, 4, 12, 0, 1.0, bias)


def _gi_estimate(gp, gn, res):
    n = gn[:, :3] / np.linalg.norm(gn[:, :3], axis=1, keepdims=True)
    to = res[0][:, :3].astype(np.float64) - gp[:, :3]
    wi = to / np.linalg.norm(to, axis=1, keepdims=True)
    cos = np.maximum((n * wi).sum(1), 0)
    return RHO / np.pi * cos * res[1][:, 0] * res[1][:, 3]


@pytest.mark.parametrize("M", [1, 8])
def test_gi_furnace_every_reservoir_estimates_rho_le(M):
    """ReSTIR GI takes its source pdf from the hit position (cos / pi toward it), so with a
    default ray bias the furnace identity f cos Le W = rho Le still holds to rounding."""
    os_, oc, gp, gn = _setup()
    res, hp, hn = os_.rgi_initial(oc, _gi_prm(M), gp, gn)
    Ms = res[2][:, 3].view(np.uint32)
    valid = Ms > 0
    assert valid.mean() > 0.99 and (Ms[valid] <= M).all()
    assert (res[1][valid, :3] == np.float32(LE)).all()
    assert np.allclose(_gi_estimate(gp, gn, res)[valid], RHO * LE, rtol=2e-5)
    # the kept normal is the interpolated one: the walls' inward normals, facing the primary
    to = res[0][valid, :3] - gp[valid, :3]
    assert ((res[2][valid, :3] * to).sum(1) < 0).all()
    assert np.array_equal(hp, gp) and np.array_equal(hn, gn)


def test_gi_temporal_reuse_with_itself_keeps_w():
    os_, oc, gp, gn = _setup()
    prm = _gi_prm(4)
    res, hp, hn = os_.rgi_initial(oc, prm, gp, gn)
    out = os_.rgi_temporal(oc, prm, gp, gn, res, res, hp, hn)
    M0, M1 = res[2][:, 3].view(np.uint32), out[2][:, 3].view(np.uint32)
    keep = (M1 == 2 * M0) & (M0 > 0)
    under = keep & (res[1][:, 3] < 50)
    assert under.mean() > 0.5
    assert np.allclose(out[1][under, 3], res[1][under, 3], rtol=1e-5)
    assert np.allclose(out[0][under, 3], 2 * res[0][under, 3], rtol=1e-6)

"""Pins for the CPU oracle (no GPU): known answers, closed forms, brute force, statistics.

Each test checks the oracle against something it was not derived from: the
published PCG constants' outputs, numpy's libm, analytic geometry, brute-force
intersection, sampling statistics, or a reduction to a textbook case.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import oracle as O
from research_amd import _lib as L
from research_amd import core
from tests.helpers import ocam, prim_ids, primary_rays, random_rays

lib = O.lib()


# ---- RNG: Common.slang:18-28, Sampling.slang:26-30 -----------------------------------
def test_pcg_hash_known_answers():
    # values of the PCG-RXS-M-XS hash listed for this path (restatement KATs)
    kat = {0: 0x07BB2FE2, 1: 0xA8BEEA3C, 2: 0x7A7ECC88, 0xDEADBEEF: 0x67299972, 0xFFFFFFFF: 0xE62A4902}
    for k, v in kat.items():
        assert lib.orc_pcg_hash(k) == v


def test_pcg_hash_matches_numpy_formula():
    # independent restatement in numpy uint32 arithmetic
    x = np.arange(0, 1 << 20, 97, dtype=np.uint64)
    st = (x * 747796405 + 2891336453) & 0xFFFFFFFF
    w = (((st >> ((st >> 28) + 4)) ^ st) * 277803737) & 0xFFFFFFFF
    ref = ((w >> 22) ^ w) & 0xFFFFFFFF
    got = np.array([lib.orc_pcg_hash(int(v)) for v in x[:4000]], dtype=np.uint64)
    assert np.array_equal(got, ref[:4000])


def test_make_seed_and_first_jitter():
    assert lib.orc_make_seed(0, 0, 0) == 0x07BB2FE2
    assert lib.orc_make_seed(0, 1, 0) == 0x9734F28B
    assert lib.orc_make_seed(959, 539, 7) == 0x67670BE6
    s = C.c_uint32(lib.orc_make_seed(0, 0, 0))
    a, b = lib.orc_rand_float(C.byref(s)), lib.orc_rand_float(C.byref(s))
    assert abs(a - 0.190399349) < 1e-8 and abs(b - 0.499476671) < 1e-8
    # float = top 24 bits * 2^-24
    assert lib.orc_uint_to_float01(0xFFFFFFFF) == np.float32((2 ** 24 - 1) / 2 ** 24)


# ---- deterministic transcendentals ---------------------------------------------------
def test_sin_cos_exp_log_accuracy():
    x = np.linspace(0.0, 2 * np.pi, 20001).astype(np.float32)
    s = np.array([lib.orc_sinf(float(t)) for t in x])
    c = np.array([lib.orc_cosf(float(t)) for t in x])
    assert np.abs(s - np.sin(x.astype(np.float64))).max() < 2.5e-7
    assert np.abs(c - np.cos(x.astype(np.float64))).max() < 2.5e-7
    y = np.linspace(-86.0, 5.0, 20001).astype(np.float32)
    e = np.array([lib.orc_expf(float(t)) for t in y])
    ref = np.exp(y.astype(np.float64))
    assert (np.abs(e - ref) / ref).max() < 3e-7
    assert lib.orc_expf(-100.0) == 0.0
    z = np.exp(np.linspace(-80, 80, 20001)).astype(np.float32)
    g = np.array([lib.orc_logf(float(t)) for t in z])
    assert (np.abs(g - np.log(z.astype(np.float64))) / np.maximum(1.0, np.abs(np.log(z)))).max() < 3e-7
    assert lib.orc_logf(0.0) == -math.inf
    assert abs(lib.orc_logf(1e-40) - math.log(1e-40)) < 1e-4  # subnormal input


def test_config_weight_exp_accuracy():
    """orc_wexpf (the A-trous config mode's weight exp: fused reduction, degree-6 polynomial in
    fmaf) against float64 exp over the weights' range, exact at 0, zero below -87.  Dropping the
    reduction's low part (about 2e-4 near -86) or the last polynomial term (about 2e-6) fails
    the bound."""
    y = np.concatenate([np.linspace(-86.5, 0.0, 40001), -np.logspace(-8, 0, 2001)]).astype(np.float32)
    e = np.array([lib.orc_wexpf(float(t)) for t in y])
    ref = np.exp(y.astype(np.float64))
    assert (np.abs(e - ref) / ref).max() < 5e-7
    assert lib.orc_wexpf(0.0) == 1.0 and lib.orc_wexpf(-100.0) == 0.0


def test_fp16_conversion_matches_numpy():
    rng = np.random.default_rng(3)
    v = np.concatenate([rng.normal(scale=s, size=3000) for s in (1e-6, 1e-3, 1.0, 1e3, 6e4)]).astype(np.float32)
    v = np.concatenate([v, np.float32([0.0, -0.0, 65504.0, 65520.0, 1e-8, 5.96e-8, 2.98e-8, 65519.0])])
    got = np.array([lib.orc_f32_to_f16(float(t)) for t in v], np.uint16)
    assert np.array_equal(got, v.astype(np.float16).view(np.uint16))
    back = np.array([lib.orc_f16_to_f32(int(h)) for h in got[:5000]], np.float32)
    assert np.array_equal(back, got[:5000].view(np.float16).astype(np.float32))


# ---- camera: Camera.cpp:27-36,130-146 ---------------------------------------------------
def test_camera_closed_form_and_library_agreement():
    eye, center = (0.3, 0.7, 3.0), (-0.2, 0.1, -1.0)
    cam = core.camera_look_at(eye, center, 45.0, 160, 90)
    oc = O.Camera()
    lib.orc_camera_look_at(O.fa(eye), O.fa(center), 45.0, 160, 90, 0.01, 1000.0, 0, C.byref(oc))
    a = np.frombuffer(bytes(cam), np.float32)[:76]
    b = np.frombuffer(bytes(oc), np.float32)[:76]
    assert np.allclose(a, b, rtol=0, atol=2e-6)
    fwd = np.array(center) - np.array(eye)
    fwd /= np.linalg.norm(fwd)
    assert np.allclose(np.array(oc.direction), fwd, atol=1e-6)
    # inv_proj row 0/1: x / P00, y / P11 with P11 = -1/tan(fov/2) (Vulkan flip)
    ip = np.array(oc.inv_proj).reshape(4, 4)
    th = math.tan(math.radians(45.0) / 2)
    assert abs(ip[0, 0] - th * 160 / 90) < 1e-6 and abs(ip[1, 1] + th) < 1e-6
    # the ray through the image centre is the forward direction; the top row points up
    d = np.array(ip[0, 0] * 0.0), np.array(0.0)
    iv = np.array(oc.inv_view).reshape(4, 4)
    centre = iv[:3, :3] @ np.array([0.0, 0.0, -1.0])
    assert np.allclose(centre / np.linalg.norm(centre), fwd, atol=1e-6)
    top = iv[:3, :3] @ np.array([0.0, ip[1, 1] * -1.0, -1.0])
    assert top @ np.array([0, 1, 0]) > centre @ np.array([0, 1, 0])


# ---- intersection ----------------------------------------------------------------------
@pytest.fixture(scope="module")
def cornell():
    sa = core.generate_scene(L.GEN_CORNELL_REF)
    return sa, O.Scene.from_arrays(sa)


def test_cornell_reference_topology(cornell):
    sa, sc = cornell
    assert sc.tri_count == 36 and len(sa.lights) == 1
    assert sa.lights[0]["type"] == 0 and np.allclose(sa.lights[0]["position"], [0, 0.9, 0])
    em = sa.mats["emissive"]
    assert (em.max(axis=1) > 0).sum() == 1 and np.allclose(em.max(), 15.0)


def test_brute_force_equals_bvh(cornell):
    sa, sc = cornell
    cam = core.camera_look_at(sa.eye, sa.center, sa.fov, 48, 48)
    rays = np.concatenate([primary_rays(cam), random_rays(20000, -0.99, 0.99, 1)])
    hb, hv = sc.trace(rays, brute=True), sc.trace(rays)
    assert np.array_equal(hb.view(np.uint32), hv.view(np.uint32))
    assert (prim_ids(hb) != 0xFFFFFFFF).mean() > 0.7  # open front face: most rays still hit


def test_brute_force_equals_bvh_heightfield():
    sa = core.generate_scene(L.GEN_HEIGHTFIELD, seed=2, scale=0.06)
    sc = O.Scene.from_arrays(sa)
    rays = random_rays(3000, [-40, 0, -40], [40, 15, 40], 5)
    hb, hv = sc.trace(rays, brute=True), sc.trace(rays)
    assert np.array_equal(hb.view(np.uint32), hv.view(np.uint32))
    assert 0.05 < (prim_ids(hb) != 0xFFFFFFFF).mean() < 0.999


def test_hit_invariants(cornell):
    sa, sc = cornell
    tris = sc.triangles().reshape(-1, 3, 3).astype(np.float64)
    rays = random_rays(5000, -0.9, 0.9, 7)
    h = sc.trace(rays)
    pid = prim_ids(h)
    ok = pid != 0xFFFFFFFF
    t, b1, b2 = h[ok, 0].astype(np.float64), h[ok, 2].astype(np.float64), h[ok, 3].astype(np.float64)
    assert (b1 >= -1e-6).all() and (b2 >= -1e-6).all() and (b1 + b2 <= 1 + 1e-6).all()
    assert (t >= 0.001).all()
    v = tris[pid[ok]]
    p_bary = (1 - b1 - b2)[:, None] * v[:, 0] + b1[:, None] * v[:, 1] + b2[:, None] * v[:, 2]
    p_ray = rays[ok, :3].astype(np.float64) + t[:, None] * rays[ok, 4:7].astype(np.float64)
    assert np.abs(p_bary - p_ray).max() < 1e-5


def test_known_analytic_hit(cornell):
    _, sc = cornell
    # from (0, 0.5, 0.9) straight at the back wall z = -1: t = 1.9, back wall is flattened tris 0-1
    rays = np.array([[0.0, 0.5, 0.9, 0.001, 0.0, 0.0, -1.0, 1e9]], np.float32)
    h = sc.trace(rays, brute=True)
    assert prim_ids(h)[0] in (0, 1) and abs(h[0, 0] - 1.9) < 1e-6


def test_watertight_shared_edge(cornell):
    _, sc = cornell
    # rays along the diagonal y == z shared by the two left-wall triangles (flattened ids 6, 7),
    # above both boxes so nothing occludes the wall
    s = np.linspace(0.01, 0.999, 4001).astype(np.float32)
    rays = np.zeros((len(s), 8), np.float32)
    rays[:, 0], rays[:, 1], rays[:, 2] = 0.9, s, s
    rays[:, 3], rays[:, 4], rays[:, 7] = 0.001, -1.0, 1e9
    h = sc.trace(rays, brute=True)
    assert set(np.unique(prim_ids(h)).tolist()) <= {6, 7}  # never falls through the crack
    assert np.allclose(h[:, 0], 1.9, atol=1e-6)


# ---- BRDF / sampling: BRDF.slang, Sampling.slang ------------------------------------------
def test_cosine_hemisphere_chi_square():
    rng = np.random.default_rng(11)
    n = 40000
    zs = np.zeros(n)
    out, pdf = (C.c_float * 3)(), C.c_float()
    for i, (u0, u1) in enumerate(rng.random((n, 2)).astype(np.float32)):
        lib.orc_sample_cosine_hemisphere(float(u0), float(u1), out, C.byref(pdf))
        zs[i] = out[2]
        if i < 200:
            assert abs(pdf.value - out[2] / math.pi) < 1e-6
            assert abs(out[0] ** 2 + out[1] ** 2 + out[2] ** 2 - 1) < 1e-5
    # cos(theta)^2 is uniform on [0, 1] under cosine sampling
    hist, _ = np.histogram(zs ** 2, bins=20, range=(0, 1))
    chi2 = ((hist - n / 20) ** 2 / (n / 20)).sum()
    assert chi2 < 45  # 19 dof, p ~ 1e-3


def test_lambert_white_furnace_weight():
    # metallic-roughness BRDF sampled by its own sampler: E[f cos / pdf] for a white diffuse
    # surface is below 1 (energy lost to F) but the estimate must be finite and positive
    n, wo = O.fa([0, 0, 1]), O.fa([0.3, 0.1, 0.948683])
    base = O.fa([1, 1, 1])
    rng = np.random.default_rng(4)
    wi, pdf, f = (C.c_float * 3)(), C.c_float(), (C.c_float * 3)()
    acc = 0.0
    N = 20000
    for u0, u1 in rng.random((N, 2)).astype(np.float32):
        lib.orc_brdf_sample(wo, n, base, 0.0, 1.0, float(u0), float(u1), wi, C.byref(pdf), f)
        acc += f[0] * max(wi[2], 0.0) / pdf.value
    est = acc / N
    assert 0.8 < est < 1.05


# ---- accumulate: accumulate.slang:21-46 ----------------------------------------------------
def test_accumulate_running_mean():
    rng = np.random.default_rng(2)
    frames = rng.random((16, 100, 4)).astype(np.float32)
    acc = np.zeros((100, 4), np.float32)
    acc64 = np.zeros((100, 4), np.float64)
    for k, f in enumerate(frames):
        lib.orc_accumulate(O.p(f), O.p(acc), 100, 0, k)
        lib.orc_accumulate_f64(O.p(f), O.p(acc64), 100, 0, k)
    assert np.allclose(acc, frames.astype(np.float64).mean(0), atol=2e-6)
    assert np.allclose(acc64 / 16, frames.astype(np.float64).mean(0), rtol=0, atol=1e-15)
    lib.orc_accumulate(O.p(frames[0]), O.p(acc), 100, 1, 9)  # camera moved -> reset
    assert np.array_equal(acc, frames[0])


# ---- ReSTIR DI: with one light every candidate is that light, so W = w_sum / (M p_hat) = N = 1 ----
def test_di_single_light_identity(cornell):
    sa, sc = cornell
    cam = core.camera_look_at(sa.eye, sa.center, sa.fov, 32, 32)
    prm = O.DIParams(0, 16, 1, 0, 0, 15)
    gp, gn, rs, uv = sc.di_initial(ocam(cam), prm)
    hit = gp[:, 3] >= 0
    sel = rs[:, 0].view(np.uint32)
    M = rs[:, 3].view(np.uint32)
    lp = np.array([0, 0.9, 0], np.float32)
    to = lp - gp[:, :3]
    d2 = np.maximum((to * to).sum(1), 1e-6)
    ndl = np.maximum((gn[:, :3] * to).sum(1) / np.sqrt((to * to).sum(1)), 0)
    phat = ndl * (0.2126 * 15 + 0.7152 * 15 + 0.0722 * 12) / d2
    lit = hit & (phat > 1e-6)
    assert (sel[lit] == 0).all() and (M[lit] == 16).all()
    assert np.allclose(rs[lit, 2], 1.0, rtol=1e-6)
    assert np.allclose(rs[lit, 1], 16 * phat[lit], rtol=2e-5)  # w_sum = M p_hat N


# ---- A-trous: atrous.slang ------------------------------------------------------------------
def _gbuf(H, W, seed=0):
    rng = np.random.default_rng(seed)
    pos = np.zeros((H, W, 4), np.float32)
    pos[..., :3] = rng.random((H, W, 3)) * 0.1
    pos[..., 3] = 1.0
    nrm = np.zeros((H, W, 4), np.float32)
    nrm[..., 2] = 1.0
    return pos, nrm


def test_atrous_constant_image_is_fixed_point():
    H, W = 24, 40
    pos, nrm = _gbuf(H, W)
    img = np.full((H, W, 4), 0.37, np.float32)
    out = O.atrous_ref(img, pos, nrm, 2)
    assert np.allclose(out, img, rtol=0, atol=1e-6)


def test_atrous_reduces_to_b3_wavelet():
    H, W = 20, 28
    rng = np.random.default_rng(1)
    img = rng.random((H, W, 4)).astype(np.float32)
    pos, nrm = _gbuf(H, W)
    out = O.atrous_ref(img, pos, nrm, 1, 1e6, 1e6, 1e6)
    k = np.array([1, 4, 6, 4, 1], np.float64) / 16
    pad = np.pad(img.astype(np.float64), ((2, 2), (2, 2), (0, 0)), mode="edge")
    ref = np.zeros((H, W, 4))
    for dy in range(5):
        for dx in range(5):
            ref += k[dy] * k[dx] * pad[dy:dy + H, dx:dx + W]
    assert np.allclose(out[..., :3], ref[..., :3], atol=2e-6)
    assert np.array_equal(out[..., 3], img[..., 3])


# ---- vMF guiding: GuidingGrid.slang ------------------------------------------------------------
def _fib_sphere(n):
    i = np.arange(n) + 0.5
    phi = np.arccos(1 - 2 * i / n)
    th = np.pi * (1 + 5 ** 0.5) * i
    return np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)], 1).astype(np.float32)


@pytest.mark.parametrize("kappa", [0.5, 4.0, 60.0])
def test_vmf_pdf_integrates_to_one(kappa):
    dirs = _fib_sphere(200000)
    mu = O.fa([0.2, -0.3, 0.932738])
    vals = np.array([lib.orc_vmf_pdf(mu, kappa, O.fa(d)) for d in dirs[::4]])
    assert abs(vals.mean() * 4 * np.pi - 1.0) < 2e-3


def test_vmf_sampler_mean_resultant():
    # E[mu . w] = coth(k) - 1/k for a vMF lobe
    k = 8.0
    mu = O.fa([0.0, 0.6, 0.8])
    rng = np.random.default_rng(9)
    d, pdf = (C.c_float * 3)(), C.c_float()
    acc = np.zeros(3)
    N = 30000
    for u0, u1 in rng.random((N, 2)).astype(np.float32):
        lib.orc_vmf_sample(mu, k, float(u0), float(u1), d, C.byref(pdf))
        acc += np.array(d[:])
        if N < 0:
            break
    r = np.linalg.norm(acc / N)
    assert abs(r - (1 / np.tanh(k) - 1 / k)) < 0.01
    assert np.allclose(acc / np.linalg.norm(acc), [0, 0.6, 0.8], atol=0.02)
    # returned pdf equals vmf_pdf at the sample
    assert abs(pdf.value - lib.orc_vmf_pdf(mu, k, d)) < 1e-3 * pdf.value


def _vmf_draw(mu, k, n, rng):
    # Wood (1994) sampler for S^2, independent of the oracle's inverse CDF
    u = rng.random(n)
    w = 1 + np.log(u + (1 - u) * np.exp(-2 * k)) / k
    phi = 2 * np.pi * rng.random(n)
    s = np.sqrt(np.maximum(0, 1 - w * w))
    loc = np.stack([s * np.cos(phi), s * np.sin(phi), w], 1)
    mu = np.asarray(mu, np.float64) / np.linalg.norm(mu)
    a = np.array([1.0, 0, 0]) if abs(mu[0]) < 0.9 else np.array([0, 1.0, 0])
    t = np.cross(a, mu); t /= np.linalg.norm(t)
    b = np.cross(mu, t)
    return (loc[:, :1] * t + loc[:, 1:2] * b + loc[:, 2:] * mu).astype(np.float32)


def test_guiding_fit_recovers_kappa():
    # within the per-cell budget of GuidingGrid.slang:178-182 (2^31 / 2^20 = 2048 unit splats)
    rng = np.random.default_rng(5)
    for k_true in (2.0, 10.0, 50.0):
        dirs = _vmf_draw([0.3, 0.4, 0.866], k_true, 2000, rng)
        acc = np.zeros((1, 4), np.uint32)
        fit = np.zeros((1, 4), np.float32)
        for d in dirs:
            lib.orc_guiding_splat(O.p(acc), 0, O.fa(d))
        assert acc[0, 3] == 2000
        lib.orc_guiding_fit_cell(O.p(acc), O.p(fit), 0, 8, 512.0)
        assert (acc == 0).all()  # accumulator reset
        # Banerjee's approximation + sampling noise at n = 2000
        assert abs(fit[0, 3] - k_true) / k_true < 0.12
        assert fit[0, :3] @ np.array([0.3, 0.4, 0.866]) / np.linalg.norm([0.3, 0.4, 0.866]) > 0.99


def test_guiding_splat_fixed_point_wraps_like_int32():
    # the accumulator is two's-complement int32 in 2^-20 units: 4096 splats of +x wrap to -2^31
    acc = np.zeros((1, 4), np.uint32)
    for _ in range(4096):
        lib.orc_guiding_splat(O.p(acc), 0, O.fa([1.0, 0.0, 0.0]))
    assert acc[0, 0] == (4096 * 2 ** 20) % 2 ** 32 and acc[0, 3] == 4096
    acc2 = np.zeros((1, 4), np.uint32)
    lib.orc_guiding_splat(O.p(acc2), 0, O.fa([0.5 / 2 ** 20, -1.5 / 2 ** 20, 2.5 / 2 ** 20]))
    assert acc2[0, :3].view(np.int32).tolist() == [1, -2, 3]  # round half away from zero


def test_guiding_cell_index_quantises_positions():
    """What fixes the cell index without the hash itself (which stays unpinned): positions in
    one cell of side base + slope x distance share an index, a cell boundary, another BSDF id
    or a larger cell (farther camera) changes it, and the index stays below the cell count."""
    cam0 = O.fa([0.0, 0.0, 0.0])
    idx = lambda pos, bsdf=0, cam=cam0, base=0.1, slope=0.0, n=1 << 20: lib.orc_guiding_cell_index(
        O.fa(pos), bsdf, cam, base, slope, n)
    rng = np.random.default_rng(5)
    for _ in range(200):
        q = rng.integers(-50, 50, 3)                       # a cell of side 0.1
        inside = [(q + rng.uniform(0.05, 0.95, 3)) * 0.1 for _ in range(4)]
        ids = {idx(p.tolist()) for p in inside}
        assert len(ids) == 1
        # one step across each face lands elsewhere (a collision in 2^20 cells is ~1e-6)
        for k in range(3):
            e = np.zeros(3); e[k] = 0.1
            assert idx((inside[0] + e).tolist()) not in ids
        assert idx(inside[0].tolist(), bsdf=1) not in ids
    # cell side grows with distance: at slope 1 from 9.95 m away the side is 10.05, so points
    # 1 m apart share a cell that a 0.1 m grid splits
    far = O.fa([0.0, 0.0, -9.95])
    assert idx([0.2, 0.2, 0.0], cam=far, slope=1.0) == idx([1.2, 1.2, 0.0], cam=far, slope=1.0)
    assert idx([0.2, 0.2, 0.0]) != idx([1.2, 1.2, 0.0])
    assert max(idx(rng.uniform(-100, 100, 3).tolist(), n=1000) for _ in range(500)) < 1000


def test_guiding_fit_below_min_samples_marks_invalid():
    acc = np.zeros((2, 4), np.uint32)
    fit = np.ones((2, 4), np.float32)
    for _ in range(5):
        lib.orc_guiding_splat(O.p(acc), 1, O.fa([0, 0, 1]))
    lib.orc_guiding_fit_cell(O.p(acc), O.p(fit), 1, 8, 512.0)
    assert fit[1, 3] == -1.0


def test_em_single_lobe_equals_reference_fit():
    rng = np.random.default_rng(8)
    dirs = _vmf_draw([-0.5, 0.5, 0.707], 6.0, 2000, rng)  # within the 2048-splat budget
    acc = np.zeros((1, 4), np.uint32)
    fit = np.zeros((1, 4), np.float32)
    for d in dirs:
        lib.orc_guiding_splat(O.p(acc), 0, O.fa(d))
    lib.orc_guiding_fit_cell(O.p(acc), O.p(fit), 0, 1, 512.0)
    lobes = np.array([[0, 0, 1, 1.0, 1.0]], np.float32)
    stats = np.zeros((1, 1, 4), np.float64)
    pos = np.full((len(dirs), 3), 0.5, np.float32)
    w = np.ones(len(dirs), np.float32)
    lib.orc_vmm_estep(O.p(lobes), 1, O.p(pos), O.p(dirs), O.p(w), len(dirs), O.fa([0, 0, 0]), O.fa([1, 1, 1]), 1,
                      O.p(stats))
    lib.orc_vmm_mstep(O.p(stats), 1, 1, 512.0, 0.0, O.p(lobes))
    assert np.allclose(lobes[0, :3], fit[0, :3], atol=2e-5)
    assert abs(lobes[0, 3] - fit[0, 3]) / fit[0, 3] < 1e-4
    assert lobes[0, 4] == 1.0


def test_em_recovers_two_lobe_mixture_and_likelihood_increases():
    rng = np.random.default_rng(12)
    m1, m2 = np.array([0, 0, 1.0]), np.array([1.0, 0, 0])
    d = np.concatenate([_vmf_draw(m1, 20.0, 6000, rng), _vmf_draw(m2, 20.0, 2000, rng)])
    pos = np.full((len(d), 3), 0.5, np.float32)
    w = np.ones(len(d), np.float32)
    K = 2
    lobes = np.array([[0.577, 0.577, 0.577, 2.0, 0.5], [-0.577, -0.577, 0.577, 2.0, 0.5]], np.float32)
    ll_prev = -np.inf
    for it in range(25):
        ll = sum(math.log(lib.orc_vmm_pdf(O.p(lobes), K, O.fa(x))) for x in d[::8])
        assert ll >= ll_prev - 1e-6
        ll_prev = ll
        st = np.zeros((1, K, 4), np.float64)
        lib.orc_vmm_estep(O.p(lobes), K, O.p(pos), O.p(d), O.p(w), len(d), O.fa([0, 0, 0]), O.fa([1, 1, 1]), 1, O.p(st))
        lib.orc_vmm_mstep(O.p(st), 1, K, 512.0, 0.0, O.p(lobes))
    order = np.argsort(-lobes[:, 4])
    big, small = lobes[order[0]], lobes[order[1]]
    assert abs(big[4] - 0.75) < 0.02 and big[:3] @ m1 > 0.99 and small[:3] @ m2 > 0.99
    assert abs(big[3] - 20) < 3 and abs(small[3] - 20) < 4


# ---- G-buffer (GBufferPass by pixel-centre rays) ------------------------------------------
def _pixel_centre_dirs(cam):
    """Pixel-centre directions from the camera matrices (numpy, float64)."""
    W, H = cam.screen_width, cam.screen_height
    ip = np.asarray(cam.inv_proj, np.float64).reshape(4, 4)
    iv = np.asarray(cam.inv_view, np.float64).reshape(4, 4)
    px, py = np.meshgrid(np.arange(W) + 0.5, np.arange(H) + 0.5)
    ndc = np.stack([px.ravel() / W * 2 - 1, py.ravel() / H * 2 - 1, np.ones(W * H), np.ones(W * H)], 1)
    v = ndc @ ip.T
    w = np.stack([v[:, 0], v[:, 1], -np.ones(W * H), np.zeros(W * H)], 1) @ iv.T
    return w[:, :3] / np.linalg.norm(w[:, :3], axis=1, keepdims=True)


@pytest.mark.parametrize("kind,eye", [(L.GEN_CORNELL_REF, None), (L.GEN_CORNELL_REF, (0.0, 0.0, -4.0)),
                                      (L.GEN_CORNELL_RANDOM, None)])
def test_gbuffer_is_the_nearest_front_face_on_each_pixel_ray(kind, eye):
    """Every G-buffer surface lies on its pixel-centre ray, is the two-sided closest hit when that
    hit faces the ray and lies beyond it when it does not (back faces culled); material ids are
    the hit triangle's mesh material; normals are the flat walls' unit normals, w = material;
    misses keep the clear values.  The reference Cornell box's side walls and box tops are wound
    away from their normals (Scene.cpp:94-141 keeps the corner order), so they are culled from
    inside; from behind the back wall (eye at z = -4) that wall is culled too; the generated
    Cornell box faces every surface along its normal, so nothing in view is culled."""
    sa = core.generate_scene(kind, 3, 1.0)
    sc = O.Scene.from_arrays(sa)
    W, H = 40, 30
    cam = core.camera_look_at(eye or sa.eye, sa.center, sa.fov, W, H)
    oc = ocam(cam)
    pos, nrm, mat = sc.gbuffer(oc)
    d = _pixel_centre_dirs(oc)
    e = np.asarray(oc.position, np.float64)
    rays = np.zeros((W * H, 8), np.float32)
    rays[:, :3], rays[:, 3], rays[:, 4:7], rays[:, 7] = e, 0.0, d, 1e9
    two = sc.trace(rays)
    tid = prim_ids(two)
    tri = sc.triangles().reshape(-1, 3, 3).astype(np.float64)
    gn = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    # material of each flattened triangle: instances in order, each mesh's triangles in order
    tri_mat = np.concatenate([np.full(sa.meshes["index_count"][m] // 3, sa.meshes["material_index"][m])
                              for m in sa.insts["mesh_index"]])
    hit = mat != 0xFFFFFFFF
    assert hit.mean() > 0.15
    assert (pos[~hit] == 0).all() and (nrm[~hit] == 0).all()
    to = pos[hit, :3] - e
    dist = np.linalg.norm(to, axis=1)
    assert np.allclose(np.cross(to / dist[:, None], d[hit]), 0, atol=1e-5)  # on the pixel ray
    assert (pos[hit, 3] == 1).all() and (nrm[hit, 3].view(np.uint32) == mat[hit]).all()
    assert np.allclose(np.linalg.norm(nrm[hit, :3], axis=1), 1, atol=1e-6)
    any_hit = tid != 0xFFFFFFFF
    facing = np.zeros(W * H, bool)
    facing[any_hit] = (gn[tid[any_hit]] * d[any_hit]).sum(1) < 0
    # the numpy rays differ from the G-buffer's fp32 ones in the last bits: silhouette pixels
    # may disagree on hit / miss
    assert (~hit[any_hit & facing]).sum() <= 0.01 * W * H and hit[~any_hit].sum() <= 0.01 * W * H
    culled = any_hit & ~facing  # the two-sided closest hit faces away
    two_hit = hit & any_hit
    front, back = two_hit & facing, two_hit & ~facing
    full_dist = np.zeros(W * H)
    full_dist[hit] = dist
    same = front & np.isclose(full_dist, two[:, 0], rtol=1e-5)
    assert same.sum() >= 0.98 * front.sum()  # the rest: silhouette pixels of the two ray sets
    if same.any():
        assert (mat[same] == tri_mat[tid[same]]).mean() > 0.98
        n_geo = gn[tid[same]] / np.linalg.norm(gn[tid[same]], axis=1, keepdims=True)
        assert (np.abs(np.abs((nrm[same, :3] * n_geo).sum(1)) - 1) < 1e-5).mean() > 0.98
    # a back-facing two-sided hit: the G-buffer surface lies at or beyond it (at: the ray meets
    # an edge shared with a front face, or a coincident front face)
    assert (dist[back[hit]] >= two[back, 0] * (1 - 1e-5)).all()
    if back.any():
        assert (dist[back[hit]] > two[back, 0] * (1 + 1e-4)).mean() > 0.5
    if kind == L.GEN_CORNELL_REF:
        assert culled.mean() > 0.3  # culled walls: the box's far side (or nothing) shows through
    else:
        assert not culled.any() and front.mean() > 0.8

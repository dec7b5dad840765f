"""Independent pins of the oracle's metallic-roughness BSDF (BRDF.slang:16-187).

Every expectation here is computed in float64 from the published definitions, not from
the oracle's code:
  * Fresnel: Schlick 1994, F = F0 + (1 - F0)(1 - cos)^5.
  * GGX normal distribution: Walter et al. 2007 eq. 33, written in the tangent form
    D = a^2 / (pi cos^4 (a^2 + tan^2)^2), whose hemisphere integral of D cos is 1.
  * Geometry: Karis 2013's Schlick-GGX G1(x) = x / (x (1 - k) + k), k = a/2, G = G1(v) G1(l).
  * Visible normals: Heitz 2018 (JCGT 7(4)), D_wo(h) = G1_smith(wo) max(0, wo.h) D(h) / cos(wo)
    with the exact Smith G1 = 2 / (1 + sqrt(1 + a^2 tan^2)); reflection Jacobian 1 / (4 wo.h).
The roughness convention is the shader's: GGX alpha = clamp(roughness, .04, 1)^2.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from tests.bsdf64 import cook_torrance as _ct, d_ggx, g1_schlick, g1_smith_exact, normalize

F32 = np.float32


def lib():
    return O.lib()


def fa(v):
    return O.fa(v)


def sph(theta, phi):
    st = np.sin(theta)
    return np.stack([st * np.cos(phi), st * np.sin(phi), np.cos(theta)], -1)


def vndf_density(h, wo, a):
    """Heitz 2018: density of visible normals h (local frame, z = normal)."""
    cos_o = wo[2]
    return g1_smith_exact(cos_o, a) * np.maximum(0.0, h @ wo) * d_ggx(h[..., 2], a) / cos_o


# ---- Fresnel -----------------------------------------------------------------------------
def test_fresnel_schlick_closed_form():
    F0 = [0.04, 0.5, 0.95]
    out = (C.c_float * 3)()
    for c in [1.0, 0.0, 0.3, 0.7, 0.999, -0.2, 1.3]:
        lib().orc_fresnel_schlick(c, fa(F0), out)
        x = min(max(1.0 - c, 0.0), 1.0)
        want = [f + (1.0 - f) * x ** 5 for f in F0]
        assert np.allclose(list(out), want, rtol=1e-6, atol=1e-7), (c, list(out), want)
    lib().orc_fresnel_schlick(1.0, fa(F0), out)
    assert list(out) == [F32(f) for f in F0]          # normal incidence: F0 exactly
    lib().orc_fresnel_schlick(0.0, fa(F0), out)
    assert list(out) == [1.0, 1.0, 1.0]               # grazing: 1 exactly


# ---- GGX D -------------------------------------------------------------------------------
@pytest.mark.parametrize("rough", [0.3, 0.5, 0.8, 1.0])
def test_ggx_distribution_matches_walter_and_is_normalised(rough):
    a = rough * rough
    theta = (np.arange(200000) + 0.5) * (np.pi / 2) / 200000
    c = np.cos(theta)
    got = np.array([lib().orc_distribution_ggx(float(x), a) for x in c[::997]])
    assert np.allclose(got, d_ggx(c[::997], a), rtol=2e-5)
    # hemisphere integral of D(h) cos(h) = 1, by quadrature on the oracle's own values
    Dv = np.array([lib().orc_distribution_ggx(float(x), a) for x in c[::10]], np.float64)
    th = theta[::10]
    integral = 2 * np.pi * np.sum(Dv * np.cos(th) * np.sin(th)) * (np.pi / 2 / len(th))
    assert abs(integral - 1.0) < 2e-3, integral


def test_ggx_distribution_epsilon_floor():
    """The shader floors the denominator at kEpsilon = 1e-4 (BRDF.slang:29): at the smallest
    roughness (.04, alpha = .0016) the peak is a^2 / 1e-4, not 1 / (pi a^2)."""
    a = F32(0.04) * F32(0.04)
    got = lib().orc_distribution_ggx(1.0, float(a))
    assert np.isclose(got, float(a) ** 2 / 1e-4, rtol=1e-5)
    # where pi d^2 >= 1e-4 the floor is inactive and Walter's form holds
    assert np.isclose(lib().orc_distribution_ggx(0.5, float(a)), d_ggx(np.array(0.5), float(a)), rtol=1e-5)


# ---- Smith G -----------------------------------------------------------------------------
def test_smith_schlick_ggx_closed_form():
    rng = np.random.default_rng(3)
    for a in [0.0016, 0.09, 0.25, 0.64, 1.0]:
        for nv, nl in rng.uniform(0.0, 1.0, size=(40, 2)):
            got = lib().orc_geometry_smith_ggx(float(nv), float(nl), a)
            want = g1_schlick(F32(nv), a) * g1_schlick(F32(nl), a)
            assert np.isclose(got, want, rtol=2e-6, atol=1e-7), (a, nv, nl)
            assert got == lib().orc_geometry_smith_ggx(float(nl), float(nv), a)  # symmetric
        assert np.isclose(lib().orc_geometry_smith_ggx(1.0, 1.0, a), 1.0, rtol=1e-6)
        assert lib().orc_geometry_smith_ggx(0.0, 0.5, a) == 0.0


# ---- brdf_eval ---------------------------------------------------------------------------
def cook_torrance(n, wo, wi, base, metallic, rough):
    return _ct(n, wo, wi, base, np.float64(metallic), np.float64(rough))


def eval_oracle(n, wo, wis, base, m, r):
    wis = np.ascontiguousarray(wis, F32)
    out = np.zeros((len(wis), 3), F32)
    lib().orc_brdf_eval_n(fa(n), fa(wo), O.p(wis), len(wis), fa(base), m, r, O.p(out))
    return out


def upper_dirs(rng, k, min_cos=0.05):
    d = normalize(rng.normal(size=(k * 3, 3)))
    d[:, 2] = np.abs(d[:, 2])
    return d[d[:, 2] > min_cos][:k]


@pytest.mark.parametrize("m,r", [(0.0, 0.5), (0.3, 0.3), (1.0, 0.7), (0.6, 1.0)])
def test_brdf_eval_matches_cook_torrance(m, r):
    rng = np.random.default_rng(7)
    n = np.array([0.0, 0.0, 1.0])
    base = [0.8, 0.5, 0.2]
    for wo in upper_dirs(rng, 6).astype(F32):
        wis = upper_dirs(rng, 200).astype(F32)
        got = eval_oracle(n, wo, wis, base, m, r)
        want = np.array([cook_torrance(n, wo.astype(np.float64), wi.astype(np.float64), base, m, r) for wi in wis])
        assert np.allclose(got, want, rtol=5e-5, atol=1e-7), np.abs(got - want).max()


def test_brdf_eval_reciprocity():
    """The specular lobe is reciprocal; the diffuse lobe is weighted by 1 - F(n.wo) (BRDF.slang:77-78),
    so swapping wo and wi changes f by exactly (F(n.wi) - F(n.wo)) (1 - m) base / pi."""
    rng = np.random.default_rng(11)
    n = np.array([0.0, 0.0, 1.0], F32)
    base = np.array([0.9, 0.6, 0.3])
    for m in (1.0, 0.4):
        F0 = 0.04 * (1 - m) + base * m
        fres = lambda c: F0 + (1 - F0) * (1 - c) ** 5
        for _ in range(20):
            wo, wi = upper_dirs(rng, 2).astype(F32)
            a = eval_oracle(n, wo, wi[None], base, m, 0.5)[0].astype(np.float64)
            b = eval_oracle(n, wi, wo[None], base, m, 0.5)[0].astype(np.float64)
            want = (fres(float(wi[2])) - fres(float(wo[2]))) * (1 - m) * base / np.pi
            assert np.allclose(a - b, want, rtol=1e-3, atol=2e-6 * max(a.max(), 1.0)), (m, a - b, want)


# ---- VNDF sampler ------------------------------------------------------------------------
def theta_edges(a, nb):
    # equal-mass bins of D cos (tan^2 = a^2 q / (1 - q)), closed at pi/2
    q = np.linspace(0, 1, nb + 1)
    with np.errstate(divide="ignore"):
        e = np.arctan(a * np.sqrt(q / (1 - q)))
    e[-1] = np.pi / 2
    return e


def chi2_ok(counts, expected, min_exp=8.0):
    keep = expected >= min_exp
    assert keep.sum() > 20
    # pool the small bins into one
    obs = np.append(counts[keep], counts[~keep].sum())
    exp = np.append(expected[keep], expected[~keep].sum())
    if exp[-1] < min_exp:
        obs, exp = obs[:-1], exp[:-1]
    chi2 = ((obs - exp) ** 2 / exp).sum()
    dof = len(obs) - 1
    # ~ 5.5 sigma above the mean of the chi^2 law: a p-value far below 1e-6
    return chi2, dof, chi2 < dof + 5.5 * np.sqrt(2 * dof)


@pytest.mark.parametrize("rough,theta_o", [(0.5, 0.0), (0.5, 0.7), (0.8, 1.2), (1.0, 0.4)])
def test_vndf_sampler_follows_heitz_density(rough, theta_o):
    a = rough * rough
    wo = np.array([np.sin(theta_o), 0.0, np.cos(theta_o)])
    # the density itself integrates to 1 (checks the expectation used below)
    th = (np.arange(3000) + 0.5) * (np.pi / 2) / 3000
    ph = (np.arange(256) + 0.5) * 2 * np.pi / 256
    T, P = np.meshgrid(th, ph, indexing="ij")
    dens = vndf_density(sph(T, P), wo, a) * np.sin(T)
    assert abs(dens.sum() * (np.pi / 2 / 3000) * (2 * np.pi / 256) - 1.0) < 2e-3
    # histogram of the oracle's samples
    N = 300_000
    u = np.random.default_rng(5).random((N, 2), dtype=F32)
    h = np.zeros((N, 3), F32)
    lib().orc_sample_ggx_vndf_n(fa(wo), rough, O.p(u), N, O.p(h))
    h = h.astype(np.float64)
    assert np.allclose(np.linalg.norm(h, axis=1), 1.0, atol=1e-5) and (h[:, 2] >= 0).all()
    te, nph = theta_edges(a, 24), 12
    pe = np.linspace(-np.pi, np.pi, nph + 1)
    ht, hp = np.arccos(np.clip(h[:, 2], -1, 1)), np.arctan2(h[:, 1], h[:, 0])
    counts, _, _ = np.histogram2d(ht, hp, bins=[te, pe])
    expected = np.zeros_like(counts)
    for i in range(len(te) - 1):
        t = te[i] + (np.arange(48) + 0.5) * (te[i + 1] - te[i]) / 48
        for j in range(nph):
            p = pe[j] + (np.arange(48) + 0.5) * (pe[j + 1] - pe[j]) / 48
            TT, PP = np.meshgrid(t, p, indexing="ij")
            expected[i, j] = (vndf_density(sph(TT, PP), wo, a) * np.sin(TT)).sum() * \
                (te[i + 1] - te[i]) / 48 * (pe[j + 1] - pe[j]) / 48 * N
    chi2, dof, ok = chi2_ok(counts.ravel(), expected.ravel())
    assert ok, (chi2, dof)


# ---- brdf_sample -------------------------------------------------------------------------
def sample_oracle(n, wo, base, m, r, N, seed):
    u = np.random.default_rng(seed).random((N, 2), dtype=F32)
    wi, pdf, f = np.zeros((N, 3), F32), np.zeros(N, F32), np.zeros((N, 3), F32)
    lib().orc_brdf_sample_n(fa(n), fa(wo), fa(base), m, r, O.p(u), N, O.p(wi), O.p(pdf), O.p(f))
    return wi.astype(np.float64), pdf.astype(np.float64), f


@pytest.mark.parametrize("m,r,theta_o", [(0.5, 0.5, 0.7), (1.0, 0.8, 0.3), (0.0, 1.0, 1.1)])
def test_brdf_sample_distribution_and_returned_pdf(m, r, theta_o):
    """brdf_sample draws the GGX visible-normal reflection with probability p_spec and the cosine
    lobe otherwise (BRDF.slang:116-172).  The pdf it returns is the mixture with the GGX *normal*
    distribution's reflection pdf D (n.h) / (4 v.h) (:162-169), which is not the density it
    samples; that is the reference's behaviour and is kept.  Specular samples below the horizon
    come back as (f = 0, pdf = 1, wi = +y)."""
    a = r * r
    n = np.array([0.0, 0.0, 1.0])
    wo = np.array([np.sin(theta_o), 0.0, np.cos(theta_o)])
    base = np.array([0.7, 0.5, 0.3])
    F0 = 0.04 * (1 - m) + base * m
    p_spec = float(np.clip((F0 + (1 - F0) * (1 - wo[2]) ** 5).mean(), 0.1, 0.9))
    N = 400_000
    wi, pdf, f = sample_oracle(n, wo, base, m, r, N, 9)
    below = (pdf == 1.0) & (np.abs(f).max(1) == 0) & (wi[:, 1] == 1.0)

    def density(w):  # sampling density over wi in the upper hemisphere
        h = normalize(w + wo)
        spec = vndf_density(h, wo, a) / (4 * np.maximum(h @ wo, 1e-300))
        return p_spec * spec + (1 - p_spec) * np.maximum(w[..., 2], 0) / np.pi

    # mass of reflections that fall below the horizon: by quadrature over visible normals
    th = (np.arange(2000) + 0.5) * (np.pi / 2) / 2000
    ph = (np.arange(512) + 0.5) * 2 * np.pi / 512
    T, P = np.meshgrid(th, ph, indexing="ij")
    H = sph(T, P)
    refl_z = 2 * (H @ wo)[..., None] * H - wo
    w_h = vndf_density(H, wo, a) * np.sin(T) * (np.pi / 2 / 2000) * (2 * np.pi / 512)
    p_below = p_spec * w_h[refl_z[..., 2] <= 0].sum()
    nb = int(below.sum())
    assert abs(nb - p_below * N) < 6 * np.sqrt(max(p_below * N, 1.0)) + 5, (nb, p_below * N)

    up = ~below
    # histogram in (sin^2 theta, phi): equal-mass bins of the cosine lobe
    s2e = np.linspace(0, 1, 21)
    pe = np.linspace(-np.pi, np.pi, 17)
    s2 = 1 - wi[up, 2] ** 2
    counts, _, _ = np.histogram2d(s2, np.arctan2(wi[up, 1], wi[up, 0]), bins=[s2e, pe])
    expected = np.zeros_like(counts)
    k = 40
    for i in range(len(s2e) - 1):
        s = s2e[i] + (np.arange(k) + 0.5) * (s2e[i + 1] - s2e[i]) / k
        t = np.arcsin(np.sqrt(s))
        for j in range(len(pe) - 1):
            p = pe[j] + (np.arange(k) + 0.5) * (pe[j + 1] - pe[j]) / k
            TT, PP = np.meshgrid(t, p, indexing="ij")
            # d(omega) = d(sin^2)/ (2 cos) d(phi)
            w = density(sph(TT, PP)) / (2 * np.cos(TT))
            expected[i, j] = w.sum() * (s2e[i + 1] - s2e[i]) / k * (pe[j + 1] - pe[j]) / k * N
    chi2, dof, ok = chi2_ok(counts.ravel(), expected.ravel())
    assert ok, (chi2, dof)

    # the returned pdf of the kept samples: p_spec D(n.h)(n.h) / (4 v.h) + (1 - p_spec) n.wi / pi
    w = wi[up][:5000]
    h = normalize(w + wo)
    nh, vh = np.maximum(h[:, 2], 1e-6), np.maximum(h @ wo, 1e-6)
    want = p_spec * d_ggx(nh, a) * nh / np.maximum(4 * vh, 1e-6) + (1 - p_spec) * np.maximum(w[:, 2], 0) / np.pi
    assert np.allclose(pdf[up][:5000], np.maximum(want, 1e-6), rtol=3e-4)


# ---- brdf_pdf ----------------------------------------------------------------------------
@pytest.mark.parametrize("r,theta_o", [(0.5, 0.0), (0.8, 0.0), (1.0, 0.0), (0.5, 1.05), (0.8, 0.6)])
def test_brdf_pdf_integrates_as_its_half_half_mixture(r, theta_o):
    """brdf_pdf (BRDF.slang:176-187) = .5 NDF reflection pdf + .5 cosine pdf over the whole
    sphere of wi: its integral is .5 (cosine lobe) + .5 x the mass of D cos over the normals h
    with v.h > 0, which is all of them at normal incidence (integral 1)."""
    a = r * r
    wo = np.array([np.sin(theta_o), 0.0, np.cos(theta_o)])
    nt, nphi = 4000, 720
    th = (np.arange(nt) + 0.5) * np.pi / nt
    ph = (np.arange(nphi) + 0.5) * 2 * np.pi / nphi
    T, P = np.meshgrid(th, ph, indexing="ij")
    wi = np.ascontiguousarray(sph(T, P).reshape(-1, 3), F32)
    out = np.zeros(len(wi), F32)
    lib().orc_brdf_pdf_n(fa([0, 0, 1]), fa(wo), O.p(wi), len(wi), r, O.p(out))
    integral = (out.reshape(nt, nphi).astype(np.float64) * np.sin(T)).sum() * (np.pi / nt) * (2 * np.pi / nphi)
    # expectation over h
    th2 = (np.arange(3000) + 0.5) * (np.pi / 2) / 3000
    T2, P2 = np.meshgrid(th2, ph, indexing="ij")
    H = sph(T2, P2)
    mass = (d_ggx(H[..., 2], a) * H[..., 2] * np.sin(T2) * (H @ wo > 0)).sum() * (np.pi / 2 / 3000) * \
        (2 * np.pi / nphi)
    want = 0.5 + 0.5 * mass
    if theta_o == 0.0:
        assert abs(mass - 1.0) < 2e-3
    assert abs(integral - want) < 3e-3, (integral, want)

"""float64 metallic-roughness BSDF from the published definitions, for checking the oracle.

Fresnel: Schlick 1994.  GGX D: Walter et al. 2007 eq. 33 (tangent form).  G: Karis 2013's
Schlick-GGX with k = alpha / 2.  Visible normals: Heitz 2018 with the exact Smith G1.
The roughness convention is the shader's (BRDF.slang:75): alpha = clamp(roughness, .04, 1)^2.
"""
from __future__ import annotations

import numpy as np


def normalize(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def d_ggx(cos_h, a):
    # Walter's a^2 / (pi cos^4 (a^2 + tan^2)^2), with cos^4 (a^2 + tan^2)^2 = (a^2 cos^2 + sin^2)^2
    c = np.clip(cos_h, 0.0, 1.0)
    d = a * a / (np.pi * (a * a * c * c + (1.0 - c * c)) ** 2)
    return np.where(cos_h > 0, d, 0.0)


def g1_smith_exact(cos_v, a):
    c = np.clip(cos_v, 1e-300, 1.0)
    tan2 = (1.0 - c * c) / (c * c)
    return 2.0 / (1.0 + np.sqrt(1.0 + a * a * tan2))


def g1_schlick(x, a):
    k = a / 2.0
    return x / (x * (1.0 - k) + k)


def fresnel(F0, cos):
    return F0 + (1.0 - F0) * np.clip(1.0 - cos, 0.0, 1.0)[..., None] ** 5


def cook_torrance(n, wo, wi, base, metallic, rough):
    """Metallic-roughness BRDF over arrays (..., 3): (1 - F(n.wo))(1 - m) base / pi
    + D(h) G(n.v, n.l) F(v.h) / max(4 n.v n.l, 1e-4).  base (..., 3), metallic / rough (...)."""
    n, wo, wi, base = (np.asarray(x, np.float64) for x in (n, wo, wi, base))
    metallic = np.asarray(metallic, np.float64)
    r = np.clip(np.asarray(rough, np.float64), 0.04, 1.0)
    a = r * r
    F0 = 0.04 * (1.0 - metallic)[..., None] + base * metallic[..., None]
    h = normalize(wo + wi)
    dot = lambda x, y: np.sum(x * y, axis=-1)
    nv, nl = np.maximum(dot(n, wo), 0.0), np.maximum(dot(n, wi), 0.0)
    spec = (d_ggx(np.maximum(dot(n, h), 0.0), a) * g1_schlick(nv, a) * g1_schlick(nl, a))[..., None] * \
        fresnel(F0, np.maximum(dot(wo, h), 0.0))
    spec = spec / np.maximum(4.0 * nv * nl, 1e-4)[..., None]
    return (1.0 - fresnel(F0, nv)) * (1.0 - metallic)[..., None] * base / np.pi + spec

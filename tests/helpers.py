"""Shared helpers for the parity tests (oracle = the checker)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from oracle import oracle as O
from research_amd import _lib as L
from research_amd import core


def ocam(cam: L.Camera) -> O.Camera:
    return O.camera_from_bytes(bytes(cam))


def primary_rays(cam: L.Camera, frame: int = 0) -> np.ndarray:
    """The reference's jittered primary rays (pathtracer.slang:192-205) from the oracle."""
    l = O.lib()
    W, H = cam.screen_width, cam.screen_height
    oc = ocam(cam)
    rays = np.zeros((W * H, 8), np.float32)
    o, d = (C.c_float * 3)(), (C.c_float * 3)()
    for py in range(H):
        for px in range(W):
            seed = C.c_uint32(l.orc_make_seed(px, py, frame))
            l.orc_primary_ray(C.byref(oc), px, py, C.byref(seed), o, d)
            rays[py * W + px] = [o[0], o[1], o[2], 0.001, d[0], d[1], d[2], 1e9]
    return rays


def random_rays(n: int, lo, hi, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, size=(n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = o
    rays[:, 3] = 0.001
    rays[:, 4:7] = d
    rays[:, 7] = 1e9
    return rays


def prim_ids(h: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(h[:, 1]).view(np.uint32)


class restir_cull:
    """ReSTIR PT / GI rays cull back faces in both the library and the oracle while active
    (default: two-sided, as the reference's CULL_DISABLE instances trace)."""

    def __init__(self, ctx, on: int):
        self.ctx, self.on = ctx, int(on)

    def __enter__(self):
        L.call("rrh_ctx_set_option", self.ctx.h, L.OPT_RESTIR_CULL, self.on)
        O.lib().orc_set_restir_cull(self.on)
        return self

    def __exit__(self, *exc):
        L.call("rrh_ctx_set_option", self.ctx.h, L.OPT_RESTIR_CULL, 0)
        O.lib().orc_set_restir_cull(0)
        return False


def cornell_sun_mr():
    """The reference Cornell box with metallic-roughness materials and, beside its point light, a
    directional light entering through the open front: the path tracer's directional NEE uses
    brdf_pdf for its MIS weight (pathtracer.slang:146-150,171-173) and ReSTIR DI takes its
    directional branches (restir_di.slang:169-173, 444-448)."""
    sa = core.generate_scene(L.GEN_CORNELL_REF)
    mats = sa.mats.copy()
    mats["metallic"] = [0.0, 0.7, 0.3, 0.0]
    mats["roughness"] = [0.6, 0.25, 0.45, 1.0]
    sun = np.zeros(1, L.LIGHT)
    sun["type"], sun["direction"], sun["emission"], sun["intensity"] = 1, (-0.25, -0.35, -1.0), (3.0, 3.0, 2.5), 1.0
    sa.mats = mats
    sa.lights = np.concatenate([sa.lights, sun])
    assert not (mats["flags"] & L.MAT_PURE_LAMBERT).any()
    return sa


def scene_arrays(kind, seed=1, scale=1.0):
    """A generated scene, or a test scene when `kind` is a function returning SceneArrays."""
    return kind() if callable(kind) else core.generate_scene(kind, seed, scale)

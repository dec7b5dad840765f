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

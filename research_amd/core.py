"""Thin object layer over the C ABI: device context, scenes, images, grids.

Memory: device buffers belong to the objects below and are released by close()
(or garbage collection).  Host data crosses the boundary only on upload/download.
"""
from __future__ import annotations

import atexit
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L

_EXITING = False


def _mark_exit() -> None:
    # at interpreter exit objects are finalised in arbitrary order: device memory is
    # released with the process instead of through per-object destroy calls
    global _EXITING
    _EXITING = True


atexit.register(_mark_exit)


class Context:
    """One HIP device + one stream (replaces the reference's rr::rhi::Device queue)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        L.call("rrh_ctx_create", device, C.byref(h))
        self.h = h

    def synchronize(self) -> None:
        L.call("rrh_ctx_synchronize", self.h)

    def device_info(self) -> dict:
        cu, clk = C.c_int(), C.c_int()
        name = C.create_string_buffer(64)
        L.call("rrh_device_info", self.h, C.byref(cu), C.byref(clk), name, 64)
        return {"cu_count": cu.value, "clock_khz": clk.value, "arch": name.value.decode()}

    def marker(self, slot: int) -> None:
        L.call("rrh_marker_record", self.h, slot)

    def elapsed_ms(self, a: int, b: int) -> float:
        ms = C.c_float()
        L.call("rrh_marker_elapsed", self.h, a, b, C.byref(ms))
        return ms.value

    def close(self) -> None:
        if getattr(self, "h", None):
            L.load().rrh_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        if _EXITING:
            return
        try:
            self.close()
        except Exception:
            pass


def camera_look_at(eye, center, fov_y_deg: float, width: int, height: int, near: float = 0.01,
                   far: float = 1000.0, frame_index: int = 0) -> L.Camera:
    cam = L.Camera()
    L.call("rrh_camera_look_at", C.byref(L.f3(eye)), C.byref(L.f3(center)), float(fov_y_deg), width, height,
           float(near), float(far), frame_index, C.byref(cam))
    return cam


def camera_bytes(cam: L.Camera) -> np.ndarray:
    return np.frombuffer(bytes(cam), dtype=np.uint8).copy()


@dataclass
class SceneArrays:
    verts: np.ndarray
    idx: np.ndarray
    meshes: np.ndarray
    mats: np.ndarray
    insts: np.ndarray
    lights: np.ndarray
    flags: int
    eye: tuple
    center: tuple
    fov: float

    @property
    def tri_count(self) -> int:
        return int(sum(self.meshes["index_count"][self.insts["mesh_index"]]) // 3)


def generate_scene(kind: int, seed: int = 1, scale: float = 1.0) -> SceneArrays:
    d = L.SceneDesc()
    L.call("rrh_scene_generate", kind, seed, float(scale), C.byref(d))
    try:
        def arr(p, n, dt):
            if n == 0:
                return np.zeros(0, dtype=dt)
            buf = (C.c_char * (n * np.dtype(dt).itemsize)).from_address(p)
            return np.frombuffer(buf, dtype=dt).copy()

        sa = SceneArrays(arr(d.verts, d.nv, L.VERTEX), arr(d.idx, d.ni, np.uint32), arr(d.meshes, d.nm, L.MESH),
                         arr(d.mats, d.nmat, L.MATERIAL), arr(d.insts, d.ninst, L.INSTANCE),
                         arr(d.lights, d.nl, L.LIGHT), int(d.flags), tuple(d.cam_eye), tuple(d.cam_center),
                         float(d.fov_y_deg))
    finally:
        L.load().rrh_scene_desc_free(C.byref(d))
    return sa


class Scene:
    """Uploaded scene: flattened triangles + BVH + shading records on the device."""

    def __init__(self, ctx: Context, sa: SceneArrays, flags: int | None = None):
        self.ctx = ctx
        self.arrays = sa
        self.flags = sa.flags if flags is None else flags
        h = C.c_void_p()
        a = sa
        L.call("rrh_scene_upload", ctx.h, L.ptr(a.verts), len(a.verts), L.ptr(a.idx), len(a.idx), L.ptr(a.meshes),
               len(a.meshes), L.ptr(a.mats), len(a.mats), L.ptr(a.insts), len(a.insts),
               L.ptr(a.lights) if len(a.lights) else None, len(a.lights), self.flags, C.byref(h))
        self.h = h
        self.camera: L.Camera | None = None

    def info(self) -> dict:
        i = L.SceneInfo()
        L.call("rrh_scene_info_get", self.h, C.byref(i))
        return {k: (list(getattr(i, k)) if k.startswith("bounds") else getattr(i, k)) for k, _ in i._fields_}

    def triangles(self) -> np.ndarray:
        n = self.info()["tri_count"]
        out = np.zeros((n, 9), dtype=np.float32)
        L.call("rrh_scene_triangles", self.h, L.ptr(out), n)
        return out

    def update_instances(self, insts: np.ndarray, mode: int = L.UPDATE_REFIT) -> None:
        """New instance transforms (same count and meshes); refit or rebuild the BVH on the device."""
        insts = np.ascontiguousarray(insts, dtype=L.INSTANCE)
        L.call("rrh_scene_update_instances", self.ctx.h, self.h, L.ptr(insts), len(insts), mode)

    def set_camera(self, cam: L.Camera) -> None:
        self.camera = cam
        L.call("rrh_camera_set", self.h, C.byref(cam))

    def close(self) -> None:
        if getattr(self, "h", None):
            L.load().rrh_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        if _EXITING:
            return
        try:
            self.close()
        except Exception:
            pass


class Image:
    def __init__(self, ctx: Context, width: int, height: int, fmt: int = L.FMT_RGBA32F):
        self.ctx, self.w, self.h_, self.fmt = ctx, width, height, fmt
        h = C.c_void_p()
        L.call("rrh_image_create", ctx.h, width, height, fmt, C.byref(h))
        self.h = h

    @property
    def nbytes(self) -> int:
        return self.w * self.h_ * L.FMT_BYTES[self.fmt]

    def _np(self) -> tuple:
        if self.fmt == L.FMT_RGBA32F:
            return np.float32, (self.h_, self.w, 4)
        if self.fmt == L.FMT_RGBA16F:
            return np.uint16, (self.h_, self.w, 4)
        if self.fmt == L.FMT_RGBA64F:
            return np.float64, (self.h_, self.w, 4)
        if self.fmt == L.FMT_RGBA8:
            return np.uint8, (self.h_, self.w, 4)
        return np.uint32, (self.h_, self.w)

    def download(self) -> np.ndarray:
        dt, shape = self._np()
        out = np.zeros(shape, dtype=dt)
        L.call("rrh_image_download", self.ctx.h, self.h, L.ptr(out), out.nbytes)
        return out

    def upload(self, a: np.ndarray) -> None:
        dt, shape = self._np()
        a = np.ascontiguousarray(a, dtype=dt).reshape(shape)
        L.call("rrh_image_upload", self.ctx.h, self.h, L.ptr(a), a.nbytes)

    def clear(self) -> None:
        L.call("rrh_image_clear", self.ctx.h, self.h)

    def close(self) -> None:
        if getattr(self, "h", None):
            L.load().rrh_image_destroy(self.h)
            self.h = None

    def __del__(self):
        if _EXITING:
            return
        try:
            self.close()
        except Exception:
            pass


def planes(images) -> C.Array:
    """Reservoir planes as the C array of image handles the rrh_rpt_* (four planes) and
    rrh_rgi_* (three planes) entry points take."""
    images = list(images)
    if len(images) not in (3, 4):
        raise ValueError("a ReSTIR PT reservoir has four planes, a ReSTIR GI reservoir three")
    return (C.c_void_p * len(images))(*[im.h.value for im in images])


def trace_closest(ctx: Context, scene: Scene, rays: np.ndarray) -> np.ndarray:
    """rays: (n, 8) float32 = origin, tmin, dir, tmax.  Returns (n, 4) float32 hits (t, prim bits, b1, b2)."""
    rays = np.ascontiguousarray(rays, dtype=np.float32)
    hits = np.zeros((len(rays), 4), dtype=np.float32)
    L.call("rrh_trace_closest", ctx.h, scene.h, L.ptr(rays), len(rays), L.ptr(hits))
    return hits


def trace_occluded(ctx: Context, scene: Scene, rays: np.ndarray) -> np.ndarray:
    rays = np.ascontiguousarray(rays, dtype=np.float32)
    occ = np.zeros(len(rays), dtype=np.uint32)
    L.call("rrh_trace_occluded", ctx.h, scene.h, L.ptr(rays), len(rays), L.ptr(occ))
    return occ


class Grid:
    def __init__(self, ctx: Context, params: L.GridParams):
        self.ctx, self.p = ctx, params
        h = C.c_void_p()
        L.call("rrh_grid_create", ctx.h, C.byref(params), C.byref(h))
        self.h = h
        if params.kind == L.GRID_HASH:
            self.ncell = params.cell_count
        else:
            self.ncell = params.res ** 3

    def buffers_shape(self):
        if self.p.kind == L.GRID_HASH:
            return (self.ncell, 4), np.float32, (self.ncell, 4), np.uint32
        K = self.p.lobes
        return (self.ncell, K, 5), np.float32, (self.ncell, K, 4), np.float64

    def download(self):
        s0, d0, s1, d1 = self.buffers_shape()
        a, b = np.zeros(s0, d0), np.zeros(s1, d1)
        L.call("rrh_grid_download", self.ctx.h, self.h, L.ptr(a), a.nbytes, L.ptr(b), b.nbytes)
        return a, b

    def upload(self, a: np.ndarray) -> None:
        s0, d0, _, _ = self.buffers_shape()
        a = np.ascontiguousarray(a, dtype=d0).reshape(s0)
        L.call("rrh_grid_upload", self.ctx.h, self.h, L.ptr(a), a.nbytes)

    def reset(self) -> None:
        L.call("rrh_grid_reset", self.ctx.h, self.h)

    def close(self) -> None:
        if getattr(self, "h", None):
            L.load().rrh_grid_destroy(self.h)
            self.h = None

    def __del__(self):
        if _EXITING:
            return
        try:
            self.close()
        except Exception:
            pass


class Mlp:
    """Learned guide: 78 -> 64x4 -> 40 BF16 MLP producing an 8-lobe vMF mixture."""

    NPARAM = L.MLP_PARAMS

    def __init__(self, ctx: Context, bounds_min, bounds_max, lr: float = 1e-3, beta1: float = 0.9,
                 beta2: float = 0.999, eps: float = 1e-8, seed: int = 1, loss_scale: float = 1024.0):
        self.ctx = ctx
        self.p = L.MlpParams(L.f3(bounds_min), L.f3(bounds_max), lr, beta1, beta2, eps, seed, loss_scale)
        h = C.c_void_p()
        L.call("rrh_mlp_create", ctx.h, C.byref(self.p), C.byref(h))
        self.h = h

    def get_weights(self) -> np.ndarray:
        w = np.zeros(self.NPARAM, np.float32)
        L.call("rrh_mlp_get_weights", self.ctx.h, self.h, L.ptr(w), self.NPARAM)
        return w

    def set_weights(self, w: np.ndarray) -> None:
        w = np.ascontiguousarray(w, dtype=np.float32)
        L.call("rrh_mlp_set_weights", self.ctx.h, self.h, L.ptr(w), len(w))

    @staticmethod
    def _v3(a) -> np.ndarray:
        return np.ascontiguousarray(a, dtype=np.float32).reshape(-1, 3)

    def infer(self, pos, nrm, wo) -> np.ndarray:
        pos, nrm, wo = self._v3(pos), self._v3(nrm), self._v3(wo)
        out = np.zeros((len(pos), 8, 5), np.float32)
        L.call("rrh_mlp_infer", self.ctx.h, self.h, L.ptr(pos), L.ptr(nrm), L.ptr(wo), len(pos), L.ptr(out))
        return out

    def loss_grad(self, pos, nrm, wo, dirs, w):
        pos, nrm, wo, dirs = self._v3(pos), self._v3(nrm), self._v3(wo), self._v3(dirs)
        w = np.ascontiguousarray(w, dtype=np.float32)
        loss, g = C.c_float(), np.zeros(self.NPARAM, np.float32)
        L.call("rrh_mlp_loss_grad", self.ctx.h, self.h, L.ptr(pos), L.ptr(nrm), L.ptr(wo), L.ptr(dirs), L.ptr(w),
               len(w), C.byref(loss), L.ptr(g))
        return loss.value, g

    def train_step(self, pos, nrm, wo, dirs, w) -> float:
        pos, nrm, wo, dirs = self._v3(pos), self._v3(nrm), self._v3(wo), self._v3(dirs)
        w = np.ascontiguousarray(w, dtype=np.float32)
        loss = C.c_float()
        L.call("rrh_mlp_train_step", self.ctx.h, self.h, L.ptr(pos), L.ptr(nrm), L.ptr(wo), L.ptr(dirs), L.ptr(w),
               len(w), C.byref(loss))
        return loss.value

    def train_from_vertices(self, batch: int) -> float:
        loss = C.c_float()
        L.call("rrh_mlp_train_from_vertices", self.ctx.h, self.h, batch, C.byref(loss))
        return loss.value

    def adam_apply(self, grad: np.ndarray) -> None:
        grad = np.ascontiguousarray(grad, dtype=np.float32)
        L.call("rrh_mlp_adam_apply", self.ctx.h, self.h, L.ptr(grad), len(grad))

    def state(self) -> tuple[int, float]:
        st, sc = C.c_uint32(), C.c_float()
        L.call("rrh_mlp_state", self.ctx.h, self.h, C.byref(st), C.byref(sc))
        return st.value, sc.value

    def close(self) -> None:
        if getattr(self, "h", None):
            L.load().rrh_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        if _EXITING:
            return
        try:
            self.close()
        except Exception:
            pass

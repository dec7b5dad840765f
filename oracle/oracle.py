"""ctypes wrapper of liboracle.so -- the CPU restatement of the reference renderer.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

_lib = None


def build() -> None:
    r = subprocess.run(["make", "-s", "-C", str(HERE)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")


class Camera(C.Structure):
    _fields_ = [("view", C.c_float * 16), ("inv_view", C.c_float * 16), ("inv_proj", C.c_float * 16),
                ("prev_view_proj", C.c_float * 16), ("position", C.c_float * 3), ("near_plane", C.c_float),
                ("direction", C.c_float * 3), ("far_plane", C.c_float), ("frame_index", C.c_uint32),
                ("screen_width", C.c_uint32), ("screen_height", C.c_uint32), ("_pad", C.c_uint32)]


class Hit(C.Structure):
    _fields_ = [("t", C.c_float), ("prim", C.c_uint32), ("b1", C.c_float), ("b2", C.c_float)]


class PTParams(C.Structure):
    _fields_ = [("max_bounces", C.c_uint32), ("frame_seed", C.c_uint32), ("rr_start", C.c_uint32),
                ("guide_mode", C.c_uint32), ("guide_alpha", C.c_float), ("_pad", C.c_uint32 * 3)]


class RayCounts(C.Structure):
    _fields_ = [("primary", C.c_uint64), ("extension", C.c_uint64), ("shadow", C.c_uint64)]


class DIParams(C.Structure):
    _fields_ = [("frame_seed", C.c_uint32), ("num_initial_candidates", C.c_uint32), ("temporal_enable", C.c_uint32),
                ("spatial_enable", C.c_uint32), ("spatial_num_neighbors", C.c_uint32), ("spatial_radius", C.c_uint32),
                ("_pad0", C.c_uint32), ("_pad1", C.c_uint32)]


class RPTParams(C.Structure):
    _fields_ = [("frame_seed", C.c_uint32), ("num_initial_candidates", C.c_uint32), ("max_bounces", C.c_uint32),
                ("temporal_enable", C.c_uint32), ("spatial_enable", C.c_uint32),
                ("spatial_num_neighbors", C.c_uint32), ("spatial_radius", C.c_uint32),
                ("include_direct_lighting", C.c_uint32), ("indirect_strength", C.c_float), ("ray_bias", C.c_float),
                ("guiding_enable", C.c_uint32), ("guiding_alpha", C.c_float)]


class RGIParams(C.Structure):
    _fields_ = [("frame_seed", C.c_uint32), ("num_initial_candidates", C.c_uint32), ("temporal_enable", C.c_uint32),
                ("spatial_enable", C.c_uint32), ("spatial_num_neighbors", C.c_uint32),
                ("spatial_radius", C.c_uint32), ("include_direct_lighting", C.c_uint32),
                ("indirect_strength", C.c_float), ("ray_bias", C.c_float), ("_pad", C.c_uint32 * 3)]


class AtrousParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("stride", C.c_uint32), ("mode", C.c_uint32),
                ("sigma_position", C.c_float), ("sigma_normal", C.c_float), ("sigma_luminance", C.c_float),
                ("_pad", C.c_float)]


P = C.c_void_p
F = C.c_float
U = C.c_uint32
FP = C.POINTER(C.c_float)
UP = C.POINTER(C.c_uint32)

_SIGS = {
    "orc_sinf": (F, [F]), "orc_cosf": (F, [F]), "orc_expf": (F, [F]), "orc_logf": (F, [F]), "orc_wexpf": (F, [F]),
    "orc_pcg_hash": (U, [U]), "orc_uint_to_float01": (F, [U]), "orc_rand_float": (F, [UP]),
    "orc_make_seed": (U, [U, U, U]),
    "orc_camera_look_at": (None, [FP, FP, F, U, U, F, F, U, C.POINTER(Camera)]),
    "orc_primary_ray": (None, [C.POINTER(Camera), U, U, UP, FP, FP]),
    "orc_scene_create": (P, [P, U, P, U, P, U, P, U, P, U, P, U, U]),
    "orc_scene_destroy": (None, [P]),
    "orc_scene_tri_count": (U, [P]),
    "orc_scene_light_count": (U, [P]),
    "orc_scene_triangles": (None, [P, P]),
    "orc_trace_brute": (None, [P, P, U, P]),
    "orc_trace_bvh": (None, [P, P, U, P]),
    "orc_occluded": (C.c_int, [P, FP, FP, F, F]),
    "orc_gbuffer": (None, [P, C.POINTER(Camera), P, P, P]),
    "orc_tonemap": (None, [P, C.c_size_t, F, P]),
    "orc_brdf_eval": (None, [FP, FP, FP, FP, F, F, FP]),
    "orc_brdf_sample": (None, [FP, FP, FP, F, F, F, F, FP, FP, FP]),
    "orc_brdf_pdf": (F, [FP, FP, FP, F]),
    "orc_sample_cosine_hemisphere": (None, [F, F, FP, FP]),
    "orc_build_onb": (None, [FP, FP, FP]),
    "orc_fresnel_schlick": (None, [F, FP, FP]),
    "orc_distribution_ggx": (F, [F, F]),
    "orc_geometry_smith_ggx": (F, [F, F, F]),
    "orc_sample_ggx_vndf_n": (None, [FP, F, P, U, P]),
    "orc_brdf_eval_n": (None, [FP, FP, P, U, FP, F, F, P]),
    "orc_brdf_pdf_n": (None, [FP, FP, P, U, F, P]),
    "orc_brdf_sample_n": (None, [FP, FP, FP, F, F, P, U, P, P, P]),
    "orc_pt_render": (None, [P, C.POINTER(Camera), C.POINTER(PTParams), P, U, P, C.POINTER(RayCounts), C.c_int]),
    "orc_accumulate": (None, [P, P, C.c_size_t, U, U]),
    "orc_accumulate_f64": (None, [P, P, C.c_size_t, U, U]),
    "orc_di_initial": (None, [P, C.POINTER(Camera), C.POINTER(DIParams), P, P, P, P, C.c_int]),
    "orc_di_temporal": (None, [P, C.POINTER(Camera), C.POINTER(DIParams), P, P, P, P, P, P, C.c_int]),
    "orc_di_spatial_shade": (None, [P, C.POINTER(Camera), C.POINTER(DIParams), P, P, P, P, P, C.c_int]),
    "orc_di_last_shadow_rays": (C.c_uint64, []),
    "orc_rpt_initial": (None, [P, C.POINTER(Camera), C.POINTER(RPTParams), P, F, F, U, P, P, P, P, P, C.c_int]),
    "orc_rpt_temporal": (None, [P, C.POINTER(Camera), C.POINTER(RPTParams), P, P, P, P, P, P, C.c_int]),
    "orc_rpt_spatial_shade": (None, [P, C.POINTER(Camera), C.POINTER(RPTParams), P, F, F, U, P, P, P, P, P,
                                     C.c_int]),
    "orc_set_restir_cull": (None, [C.c_int]),
    "orc_rgi_initial": (None, [P, C.POINTER(Camera), C.POINTER(RGIParams), P, P, P, P, P, C.c_int]),
    "orc_rgi_temporal": (None, [P, C.POINTER(Camera), C.POINTER(RGIParams), P, P, P, P, P, P, C.c_int]),
    "orc_rgi_spatial_shade": (None, [P, C.POINTER(Camera), C.POINTER(RGIParams), P, P, P, P, P, C.c_int]),
    "orc_atrous_ref": (None, [C.POINTER(AtrousParams), P, P, P, P]),
    "orc_atrous_cfg": (None, [C.POINTER(AtrousParams), P, P, P]),
    "orc_atrous_prepare": (None, [U, U, FP, FP, P, P, P, P, P]),
    "orc_f32_to_f16": (C.c_uint16, [F]),
    "orc_f16_to_f32": (F, [C.c_uint16]),
    "orc_guiding_cell_index": (U, [FP, U, FP, F, F, U]),
    "orc_guiding_splat": (None, [P, U, FP]),
    "orc_guiding_fit_cell": (None, [P, P, U, U, F]),
    "orc_vmf_sample": (None, [FP, F, F, F, FP, FP]),
    "orc_vmf_pdf": (F, [FP, F, FP]),
    "orc_sample_primary_dir": (None, [P, FP, FP, U, FP, F, F, U, U, F, UP, FP, FP]),
    "orc_vmm_estep": (None, [P, U, P, P, P, U, FP, FP, U, P]),
    "orc_vmm_mstep": (None, [P, U, U, F, C.c_double, P]),
    "orc_vmm_pdf": (C.c_double, [P, U, FP]),
    "orc_mlp_param_count": (U, []),
    "orc_mlp_encode": (None, [FP, FP, FP, FP, FP, P]),
    "orc_mlp_forward": (None, [P, P, P, P]),
    "orc_mlp_head": (None, [P, P]),
    "orc_mlp_pdf": (C.c_double, [P, FP]),
    "orc_mlp_loss_grad": (C.c_double, [P, P, P, P, U, P]),
    "orc_adam": (None, [P, P, P, P, U, U, C.c_double, C.c_double, C.c_double, C.c_double]),
}


def tonemap(hdr: np.ndarray, exposure: float = 1.0) -> np.ndarray:
    """Screenshot 8-bit values of an (..., 4) float32 image, Application.cpp:853-866 as written."""
    hdr = np.ascontiguousarray(hdr, np.float32)
    out = np.zeros(hdr.shape, np.uint8)
    lib().orc_tonemap(p(hdr), hdr.size // 4, float(exposure), p(out))
    return out


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        l = C.CDLL(str(LIB))
        for n, (res, args) in _SIGS.items():
            f = getattr(l, n)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def fa(v) -> C.Array:
    v = list(v)
    return (C.c_float * len(v))(*[float(x) for x in v])


def p(a: np.ndarray) -> C.c_void_p:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)  # the pointer object keeps `a` alive (temporaries too)


def camera_from_bytes(b: bytes) -> Camera:
    return Camera.from_buffer_copy(b)


class Scene:
    """Oracle scene over the same record arrays the library uploads."""

    def __init__(self, verts, idx, meshes, mats, insts, lights, flags: int):
        self._keep = [np.ascontiguousarray(a) for a in (verts, idx, meshes, mats, insts, lights)]
        v, i, m, mt, ins, l = self._keep
        self.h = lib().orc_scene_create(p(v), len(v), p(i), len(i), p(m), len(m), p(mt), len(mt), p(ins), len(ins),
                                        p(l) if len(l) else None, len(l), flags)

    @classmethod
    def from_arrays(cls, sa, flags: int | None = None) -> "Scene":
        return cls(sa.verts, sa.idx, sa.meshes, sa.mats, sa.insts, sa.lights, sa.flags if flags is None else flags)

    @property
    def tri_count(self) -> int:
        return lib().orc_scene_tri_count(self.h)

    @property
    def light_count(self) -> int:
        return lib().orc_scene_light_count(self.h)

    def triangles(self) -> np.ndarray:
        out = np.zeros((self.tri_count, 9), np.float32)
        lib().orc_scene_triangles(self.h, p(out))
        return out

    def trace(self, rays: np.ndarray, brute: bool = False) -> np.ndarray:
        rays = np.ascontiguousarray(rays, np.float32)
        hits = np.zeros((len(rays), 4), np.float32)
        (lib().orc_trace_brute if brute else lib().orc_trace_bvh)(self.h, p(rays), len(rays), p(hits))
        return hits

    def gbuffer(self, cam: Camera):
        """(position RGBA, normal RGBA, material id) of the raster G-buffer, by pixel-centre rays."""
        n = cam.screen_width * cam.screen_height
        pos, nrm = np.zeros((n, 4), np.float32), np.zeros((n, 4), np.float32)
        mat = np.zeros(n, np.uint32)
        lib().orc_gbuffer(self.h, C.byref(cam), p(pos), p(nrm), p(mat))
        return pos, nrm, mat

    def pt_render(self, cam: Camera, max_bounces: int, frame_seed: int, pixels: np.ndarray, rr_start: int = 3,
                  threads: int = 0):
        pixels = np.ascontiguousarray(pixels, np.uint32)
        out = np.zeros((len(pixels), 4), np.float32)
        pr = PTParams(max_bounces, frame_seed, rr_start, 0, 0.0)
        rc = RayCounts()
        lib().orc_pt_render(self.h, C.byref(cam), C.byref(pr), p(pixels), len(pixels), p(out), C.byref(rc), threads)
        return out, (rc.primary, rc.extension, rc.shadow)

    def di_initial(self, cam: Camera, prm: DIParams, threads: int = 0):
        n = cam.screen_width * cam.screen_height
        gp, gn, rs = (np.zeros((n, 4), np.float32) for _ in range(3))
        uv = np.zeros(n, np.uint32)
        lib().orc_di_initial(self.h, C.byref(cam), C.byref(prm), p(gp), p(gn), p(rs), p(uv), threads)
        return gp, gn, rs, uv

    def di_temporal(self, cam, prm, gp, gn, rc, rc_uv, rp, rp_uv, threads: int = 0):
        rc, rc_uv = rc.copy(), rc_uv.copy()
        lib().orc_di_temporal(self.h, C.byref(cam), C.byref(prm), p(gp), p(gn), p(rc), p(rc_uv), p(rp), p(rp_uv),
                              threads)
        return rc, rc_uv

    def di_spatial_shade(self, cam, prm, gp, gn, rs, uv, threads: int = 0):
        out = np.zeros((cam.screen_width * cam.screen_height, 4), np.float32)
        lib().orc_di_spatial_shade(self.h, C.byref(cam), C.byref(prm), p(gp), p(gn), p(rs), p(uv), p(out), threads)
        return out

    # ReSTIR PT (restir_pt.slang); res: (4, W*H, 4) float32 planes; grid: (fitted or accum, base
    # cell, slope, cell count) of the reference hash grid, or None
    def rpt_initial(self, cam: Camera, prm: RPTParams, gp, gn, grid=None, threads: int = 0):
        n = cam.screen_width * cam.screen_height
        res = np.zeros((4, n, 4), np.float32)
        hp, hn = np.zeros((n, 4), np.float32), np.zeros((n, 4), np.float32)
        fitted, base, slope, count = grid if grid is not None else (None, 0.0, 0.0, 1)
        lib().orc_rpt_initial(self.h, C.byref(cam), C.byref(prm), None if fitted is None else p(fitted), base, slope,
                              count, p(gp), p(gn), p(res), p(hp), p(hn), threads)
        return res, hp, hn

    def rpt_temporal(self, cam, prm, gp, gn, res, res_prev, hp_prev, hn_prev, threads: int = 0):
        res = np.ascontiguousarray(res, np.float32).copy()
        lib().orc_rpt_temporal(self.h, C.byref(cam), C.byref(prm), p(gp), p(gn), p(res),
                               p(np.ascontiguousarray(res_prev, np.float32)), p(hp_prev), p(hn_prev), threads)
        return res

    def rpt_spatial_shade(self, cam, prm, gp, gn, res, direct=None, grid=None, threads: int = 0):
        """Returns (output, accum): accum is the grid's accumulator after the splats (a copy)."""
        out = np.zeros((cam.screen_width * cam.screen_height, 4), np.float32)
        accum, base, slope, count = grid if grid is not None else (None, 0.0, 0.0, 1)
        if accum is not None:
            accum = np.ascontiguousarray(accum, np.uint32).copy()
        lib().orc_rpt_spatial_shade(self.h, C.byref(cam), C.byref(prm), None if accum is None else p(accum), base,
                                    slope, count, p(gp), p(gn), None if direct is None else p(direct),
                                    p(np.ascontiguousarray(res, np.float32)), p(out), threads)
        return out, accum

    # ReSTIR GI (restir_gi.slang); res: (3, W*H, 4) float32 planes
    def rgi_initial(self, cam: Camera, prm: RGIParams, gp, gn, threads: int = 0):
        n = cam.screen_width * cam.screen_height
        res = np.zeros((3, n, 4), np.float32)
        hp, hn = np.zeros((n, 4), np.float32), np.zeros((n, 4), np.float32)
        lib().orc_rgi_initial(self.h, C.byref(cam), C.byref(prm), p(gp), p(gn), p(res), p(hp), p(hn), threads)
        return res, hp, hn

    def rgi_temporal(self, cam, prm, gp, gn, res, res_prev, hp_prev, hn_prev, threads: int = 0):
        res = np.ascontiguousarray(res, np.float32).copy()
        lib().orc_rgi_temporal(self.h, C.byref(cam), C.byref(prm), p(gp), p(gn), p(res),
                               p(np.ascontiguousarray(res_prev, np.float32)), p(hp_prev), p(hn_prev), threads)
        return res

    def rgi_spatial_shade(self, cam, prm, gp, gn, res, direct=None, threads: int = 0):
        out = np.zeros((cam.screen_width * cam.screen_height, 4), np.float32)
        lib().orc_rgi_spatial_shade(self.h, C.byref(cam), C.byref(prm), p(gp), p(gn),
                                    None if direct is None else p(direct), p(np.ascontiguousarray(res, np.float32)),
                                    p(out), threads)
        return out

    def __del__(self):
        try:
            if self.h:
                lib().orc_scene_destroy(self.h)
        except Exception:
            pass


def atrous_ref(inp, gpos, gnrm, stride, sigma_p=0.25, sigma_n=0.2, sigma_l=1.0):
    H, W = inp.shape[:2]
    out = np.zeros_like(inp, dtype=np.float32)
    prm = AtrousParams(W, H, stride, 0, sigma_p, sigma_n, sigma_l, 0.0)
    a, b, c = (np.ascontiguousarray(v, np.float32) for v in (inp, gpos, gnrm))  # alive across the call
    lib().orc_atrous_ref(C.byref(prm), p(a), p(b), p(c), p(out))
    return out


def atrous_prepare(rad, gpos, gnrm, cam_pos, cam_dir):
    """(rad16, gbuf16) as uint16 (H, W, 4) from RGBA32F radiance and G-buffer (H, W, 4)."""
    H, W = rad.shape[:2]
    r16, g16 = np.zeros((H, W, 4), np.uint16), np.zeros((H, W, 4), np.uint16)
    a, b, c = (np.ascontiguousarray(v, np.float32) for v in (rad, gpos, gnrm))
    lib().orc_atrous_prepare(W, H, fa(cam_pos), fa(cam_dir), p(a), p(b), p(c), p(r16), p(g16))
    return r16, g16


def atrous_cfg(inp16, gbuf16, stride, sigma_z=1.0, sigma_n=128.0, sigma_l=4.0):
    H, W = inp16.shape[:2]
    out = np.zeros_like(inp16, dtype=np.uint16)
    prm = AtrousParams(W, H, stride, 1, sigma_z, sigma_n, sigma_l, 0.0)
    a, b = np.ascontiguousarray(inp16, np.uint16), np.ascontiguousarray(gbuf16, np.uint16)
    lib().orc_atrous_cfg(C.byref(prm), p(a), p(b), p(out))
    return out

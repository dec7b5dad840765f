"""ctypes binding of librrh.so (declarations in include/rrh.h).

The library is built in-tree by research_amd/build.py.  Loading fails loudly when it
is missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "librrh.so"

# ---- numpy dtypes of the scene records (GpuScene.h byte layouts) -----------------
VERTEX = np.dtype([("pos", "<f4", 3), ("normal", "<f4", 3), ("uv", "<f4", 2), ("tangent", "<f4", 4)])
MESH = np.dtype([("vertex_base", "<u4"), ("vertex_count", "<u4"), ("index_byte_offset", "<u4"),
                 ("index_count", "<u4"), ("material_index", "<u4"), ("_pad", "<u4", 3)])
MATERIAL = np.dtype([("base_color", "<f4", 4), ("metallic", "<f4"), ("roughness", "<f4"), ("emissive", "<f4", 3),
                     ("ior", "<f4"), ("albedo_tex_idx", "<u4"), ("normal_tex_idx", "<u4"),
                     ("metallic_roughness_tex_idx", "<u4"), ("emissive_tex_idx", "<u4"), ("sampler_idx", "<u4"),
                     ("flags", "<u4"), ("_pad1", "<u4"), ("_pad2", "<u4")])
INSTANCE = np.dtype([("transform", "<f4", 16), ("mesh_index", "<u4"), ("_pad", "<u4", 3)])
LIGHT = np.dtype([("position", "<f4", 3), ("type", "<u4"), ("direction", "<f4", 3), ("radius", "<f4"),
                  ("emission", "<f4", 3), ("intensity", "<f4")])
assert VERTEX.itemsize == 48 and MESH.itemsize == 32 and MATERIAL.itemsize == 72
assert INSTANCE.itemsize == 80 and LIGHT.itemsize == 48

RRH_OK = 0
MAT_PURE_LAMBERT = 1
SCENE_TRI_LIGHTS = 1
GEN_CORNELL_REF, GEN_CORNELL_RANDOM, GEN_HEIGHTFIELD, GEN_INDOOR = 0, 1, 2, 3
FMT_RGBA32F, FMT_RGBA16F, FMT_RGBA64F, FMT_R32U, FMT_RGBA8 = 0, 1, 2, 3, 4
FMT_BYTES = {FMT_RGBA32F: 16, FMT_RGBA16F: 8, FMT_RGBA64F: 32, FMT_R32U: 4, FMT_RGBA8: 4}
GUIDE_NONE, GUIDE_VMM, GUIDE_MLP = 0, 1, 2
PT_COUNT_FETCHES = 1
OPT_TRAVERSAL_STACK, OPT_XCD_REMAP, OPT_BVH_NODES, OPT_PERSISTENT, OPT_SPECULATIVE = 1, 2, 3, 4, 5
OPT_BVH_MAX_LEAF, OPT_BVH_TRAV_COST = 6, 7
OPT_BVH_BUILDER, OPT_PLOC_RADIUS, OPT_RAY_SORT, OPT_RESTIR_CULL = 8, 9, 10, 11
OPT_TRAV_TOP_NODES, OPT_TRACE_PAIRS, OPT_REFILL, OPT_SUSPEND_VISITS, OPT_RPT_STEP_VISITS = 12, 13, 14, 15, 16
TRACE_PAIRS_DEFAULT = 1  # a new context's RRH_OPT_TRACE_PAIRS
SUSPEND_VISITS_DEFAULT = 0  # a new context's RRH_OPT_SUSPEND_VISITS
BVH_BUILD_HOST, BVH_BUILD_DEVICE = 0, 1
UPDATE_REFIT, UPDATE_REBUILD = 0, 1
NODES_BINARY, NODES_WIDE4, NODES_WIDE4Q = 0, 1, 2
STACK_PRIVATE, STACK_LDS32, STACK_HYBRID16, STACK_HYBRID32, STACK_NONE = 0, 1, 2, 3, 4
GRID_HASH, GRID_VMM = 0, 1
ATROUS_REFERENCE, ATROUS_CONFIG = 0, 1
ATROUS_FAST_MATH = 1


class Camera(C.Structure):
    _fields_ = [("view", C.c_float * 16), ("inv_view", C.c_float * 16), ("inv_proj", C.c_float * 16),
                ("prev_view_proj", C.c_float * 16), ("position", C.c_float * 3), ("near_plane", C.c_float),
                ("direction", C.c_float * 3), ("far_plane", C.c_float), ("frame_index", C.c_uint32),
                ("screen_width", C.c_uint32), ("screen_height", C.c_uint32), ("_pad", C.c_uint32)]


class SceneInfo(C.Structure):
    _fields_ = [("tri_count", C.c_uint32), ("node_count", C.c_uint32), ("bvh_depth", C.c_uint32),
                ("leaf_count", C.c_uint32), ("light_count", C.c_uint32), ("tri_light_count", C.c_uint32),
                ("node4_count", C.c_uint32), ("stack4", C.c_uint32), ("device_bytes", C.c_uint64), ("build_ms", C.c_float), ("bounds_min", C.c_float * 3),
                ("bounds_max", C.c_float * 3), ("bvh_depth4", C.c_uint32)]


class SceneDesc(C.Structure):
    _fields_ = [("verts", C.c_void_p), ("nv", C.c_uint32), ("idx", C.c_void_p), ("ni", C.c_uint32),
                ("meshes", C.c_void_p), ("nm", C.c_uint32), ("mats", C.c_void_p), ("nmat", C.c_uint32),
                ("insts", C.c_void_p), ("ninst", C.c_uint32), ("lights", C.c_void_p), ("nl", C.c_uint32),
                ("flags", C.c_uint32), ("cam_eye", C.c_float * 3), ("cam_center", C.c_float * 3),
                ("fov_y_deg", C.c_float)]


class PTParams(C.Structure):
    _fields_ = [("max_bounces", C.c_uint32), ("frame_seed", C.c_uint32), ("rr_start", C.c_uint32),
                ("guide_mode", C.c_uint32), ("guide_alpha", C.c_float), ("record_vertices", C.c_uint32),
                ("flags", C.c_uint32), ("guide_bounces", C.c_uint32), ("row_begin", C.c_uint32),
                ("row_end", C.c_uint32)]


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
    _fields_ = [("iterations", C.c_uint32), ("mode", C.c_uint32), ("sigma_position", C.c_float),
                ("sigma_normal", C.c_float), ("sigma_luminance", C.c_float), ("flags", C.c_uint32),
                ("_pad", C.c_uint32 * 2)]


class GridParams(C.Structure):
    _fields_ = [("kind", C.c_int), ("cell_count", C.c_uint32), ("base_cell_size", C.c_float),
                ("distance_slope", C.c_float), ("min_samples", C.c_uint32), ("kappa_max", C.c_float),
                ("res", C.c_uint32), ("lobes", C.c_uint32), ("bounds_min", C.c_float * 3),
                ("bounds_max", C.c_float * 3), ("decay", C.c_float), ("min_weight", C.c_float)]


class MlpParams(C.Structure):
    _fields_ = [("bounds_min", C.c_float * 3), ("bounds_max", C.c_float * 3), ("lr", C.c_float),
                ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float), ("seed", C.c_uint32),
                ("loss_scale", C.c_float)]


MLP_PARAMS = 20136

P = C.c_void_p
PP = C.POINTER(C.c_void_p)
u32, i32, f32, u64, sz = C.c_uint32, C.c_int, C.c_float, C.c_uint64, C.c_size_t

# name -> (argtypes); every function returns int
SIGNATURES = {
    "rrh_ctx_create": [i32, PP],
    "rrh_ctx_destroy": [P],
    "rrh_ctx_synchronize": [P],
    "rrh_ctx_join": [P, P],
    "rrh_device_info": [P, C.POINTER(i32), C.POINTER(i32), C.c_char_p, sz],
    "rrh_marker_record": [P, i32],
    "rrh_marker_elapsed": [P, i32, i32, C.POINTER(f32)],
    "rrh_scene_upload": [P, P, u32, P, u32, P, u32, P, u32, P, u32, P, u32, u32, PP],
    "rrh_scene_destroy": [P],
    "rrh_scene_info_get": [P, C.POINTER(SceneInfo)],
    "rrh_scene_triangles": [P, P, u32],
    "rrh_scene_update_instances": [P, P, P, u32, i32],
    "rrh_gbuffer": [P, P, P, P, P],
    "rrh_tonemap": [P, P, f32, P],
    "rrh_tonemap_thresholds": [P],
    "rrh_camera_set": [P, C.POINTER(Camera)],
    "rrh_camera_look_at": [C.POINTER(f32 * 3), C.POINTER(f32 * 3), f32, u32, u32, f32, f32, u32, C.POINTER(Camera)],
    "rrh_scene_generate": [i32, u32, f32, C.POINTER(SceneDesc)],
    "rrh_scene_desc_free": [C.POINTER(SceneDesc)],
    "rrh_image_create": [P, u32, u32, i32, PP],
    "rrh_image_destroy": [P],
    "rrh_image_upload": [P, P, P, sz],
    "rrh_image_download": [P, P, P, sz],
    "rrh_image_clear": [P, P],
    "rrh_trace_closest": [P, P, P, u32, P],
    "rrh_trace_occluded": [P, P, P, u32, P],
    "rrh_trace_closest_device": [P, P, P, u32, P],
    "rrh_trace_buffers": [P, u32, PP, PP],
    "rrh_trace_buffers_free": [P, P, P],
    "rrh_copy_to_device": [P, P, P, sz],
    "rrh_copy_to_host": [P, P, P, sz],
    "rrh_pt_render": [P, P, C.POINTER(PTParams), P],
    "rrh_pt_render_noalloc": [P, P, C.POINTER(PTParams), P],
    "rrh_pt_render_guided": [P, P, C.POINTER(PTParams), P, P, P],
    "rrh_pt_ray_counts": [P, C.POINTER(RayCounts)],
    "rrh_ray_totals": [P, C.POINTER(RayCounts), i32],
    "rrh_fetch_counts": [P, C.POINTER(u64), C.POINTER(u64), C.POINTER(u64), i32],
    "rrh_fetch_counts_shadow": [P, C.POINTER(u64), C.POINTER(u64), C.POINTER(u64), i32],
    "rrh_pt_traversal_rays": [P, C.POINTER(u64), C.POINTER(u64), i32],
    "rrh_fetch_wave_max": [P, C.POINTER(u64), C.POINTER(u64), C.POINTER(u64), C.POINTER(u64), i32],
    "rrh_profile_enable": [P, i32],
    "rrh_stream_copy": [P, sz, i32, C.POINTER(f32)],
    "rrh_ctx_set_option": [P, i32, i32],
    "rrh_profile_traversal": [P, C.POINTER(f32), C.POINTER(u32), i32],
    "rrh_profile_mark": [P],
    "rrh_profile_spans": [P, P, C.POINTER(f32), C.POINTER(f32), u32, C.POINTER(u32)],
    "rrh_accumulate": [P, P, P, u32, u32],
    "rrh_accumulate_f64": [P, P, P, u32, u32],
    "rrh_accumulate_rows": [P, P, P, u32, u32, u32, u32],
    "rrh_di_initial": [P, P, C.POINTER(DIParams), P, P, P, P],
    "rrh_di_temporal": [P, P, C.POINTER(DIParams), P, P, P, P, P, P],
    "rrh_di_spatial_shade": [P, P, C.POINTER(DIParams), P, P, P, P, P],
    "rrh_di_ray_counts": [P, C.POINTER(RayCounts)],
    "rrh_rpt_initial": [P, P, C.POINTER(RPTParams), P, P, P, P, P, P],
    "rrh_rpt_temporal": [P, P, C.POINTER(RPTParams), P, P, P, P, P, P],
    "rrh_rpt_spatial_shade": [P, P, C.POINTER(RPTParams), P, P, P, P, P, P],
    "rrh_rpt_clear": [P, P, P, P],
    "rrh_rpt_ray_counts": [P, C.POINTER(RayCounts)],
    "rrh_rgi_initial": [P, P, C.POINTER(RGIParams), P, P, P, P, P],
    "rrh_rgi_temporal": [P, P, C.POINTER(RGIParams), P, P, P, P, P, P],
    "rrh_rgi_spatial_shade": [P, P, C.POINTER(RGIParams), P, P, P, P, P],
    "rrh_rgi_clear": [P, P, P, P],
    "rrh_rgi_ray_counts": [P, C.POINTER(RayCounts)],
    "rrh_atrous": [P, C.POINTER(AtrousParams), P, P, P, P, P, C.POINTER(i32)],
    "rrh_atrous_prepare": [P, P, P, P, P, P, P],
    "rrh_grid_create": [P, C.POINTER(GridParams), PP],
    "rrh_grid_destroy": [P],
    "rrh_grid_reset": [P, P],
    "rrh_grid_splat": [P, P, P, P, P, u32, C.POINTER(f32 * 3)],
    "rrh_grid_fit": [P, P],
    "rrh_grid_sample_primary": [P, P, P, P, P, P, u32, C.POINTER(f32 * 3), f32, P, P],
    "rrh_grid_em_step": [P, P, P, P, P, u32],
    "rrh_grid_em_from_vertices": [P, P],
    "rrh_grid_download": [P, P, P, sz, P, sz],
    "rrh_grid_upload": [P, P, P, sz],
    "rrh_mlp_create": [P, C.POINTER(MlpParams), PP],
    "rrh_mlp_destroy": [P],
    "rrh_mlp_set_weights": [P, P, P, u32],
    "rrh_mlp_get_weights": [P, P, P, u32],
    "rrh_mlp_infer": [P, P, P, P, P, u32, P],
    "rrh_mlp_loss_grad": [P, P, P, P, P, P, P, u32, P, P],
    "rrh_mlp_train_step": [P, P, P, P, P, P, P, u32, P],
    "rrh_mlp_train_from_vertices": [P, P, u32, P],
    "rrh_mlp_adam_apply": [P, P, P, u32],
    "rrh_mlp_state": [P, P, P, P],
}


class RRHError(RuntimeError):
    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed ({code}): {msg}")
        self.code = code


_lib = None


def load() -> C.CDLL:
    """Load librrh.so; raises if it was not built (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m research_amd.build` "
                           "(the renderer has no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH))
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.rrh_last_error.argtypes = []
    lib.rrh_last_error.restype = C.c_char_p
    _lib = lib
    return lib


def call(name: str, *args) -> None:
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != RRH_OK:
        msg = lib.rrh_last_error()
        raise RRHError(name, rc, msg.decode() if msg else "")


def ptr(a: np.ndarray | None) -> C.c_void_p | None:
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays passed to librrh must be C-contiguous"
    return a.ctypes.data_as(C.c_void_p)  # the pointer object keeps `a` alive (temporaries too)


def f3(v) -> C.Array:
    return (C.c_float * 3)(*[float(x) for x in v])


def declared_symbols(header: Path | None = None) -> list[str]:
    """Entry points declared in include/rrh.h (used by the export check)."""
    import re

    header = header or (HERE.parent / "include" / "rrh.h")
    text = header.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(rrh_\w+)\s*\(", text, re.M)))

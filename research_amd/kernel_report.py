"""Achieved bandwidth of the passes that stream images rather than walk the BVH: frame
accumulation, ReSTIR DI's temporal stage, the A-trous filter in both modes (with the fp16
input preparation; exact and with RRH_ATROUS_FAST_MATH), the tonemap, at 1920x1080 on the 1,164,288-triangle heightfield, beside a
plain device copy (the measured HBM ceiling).  Each figure is algorithmic bytes per pixel x
pixels / time per call, the time from HIP events on the library's stream around each call
after 1 GiB has been written elsewhere (so the inputs start in HBM, not in the 256 MB Infinity
Cache); the back-to-back time of `reps` calls is kept beside it as warm_ms.  Inputs come from one
path-traced + ReSTIR DI frame of the same scene.

The vMF grid's EM fit runs last on 2,097,152 synthetic samples (positions uniform in the room
scene's bounds, directions uniform, weights in (0, 1]), three E-step + M-step rounds through
rrh_grid_em_step; that entry point uploads its host arrays first, so its kernels are timed by
a kernel trace of this module, not here.

Run:  python -m research_amd.kernel_report [--reps 20]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math

import numpy as np

from . import _lib as L
from . import core

# algorithmic bytes per pixel (per call; A-trous per iteration)
BYTES = {
    "accumulate": 16 + 16 + 16,           # radiance in, running mean in, running mean out
    "di_temporal": 2 * 20 + 32 + 20,      # two reservoirs (RGBA32F + R32U) + G-buffer in, reservoir out
    "atrous_reference": 16 + 32 + 16,     # RGBA32F radiance + position + normal in, radiance out
    "atrous_prepare": 16 + 32 + 8 + 8,    # radiance + G-buffer in, fp16 radiance + fp16 normal/depth out
    "atrous_config": 8 + 8 + 8,           # fp16 radiance + fp16 G-buffer in, fp16 radiance out
    "atrous_reference_fast": 16 + 32 + 16,
    "atrous_config_fast": 8 + 8 + 8,
    "tonemap": 16 + 4,                    # RGBA32F in, RGBA8 out
}


def _timed(ctx: core.Context, reps: int, fn, flush) -> tuple:
    """(cold, warm) ms per call.  Cold: each call on its own between two markers, after `flush`
    has written 1 GiB elsewhere, so the call's working set starts outside the 256 MB Infinity
    Cache (HBM figures).  Warm: `reps` back-to-back calls, whose inputs stay cache resident."""
    fn()
    cold = 0.0
    for _ in range(reps):
        flush()
        ctx.marker(2)
        fn()
        ctx.marker(3)
        ctx.synchronize()
        cold += ctx.elapsed_ms(2, 3)
    ctx.marker(0)
    for _ in range(reps):
        fn()
    ctx.marker(1)
    ctx.synchronize()
    return cold / reps, ctx.elapsed_ms(0, 1) / reps


def report(ctx: core.Context, W: int = 1920, H: int = 1080, reps: int = 20, em_samples: int = 1 << 21) -> dict:
    c = ctx.h
    g = C.c_float()
    L.call("rrh_stream_copy", c, 1 << 30, 10, C.byref(g))
    out = {"width": W, "height": H, "reps": reps, "copy_gbs": round(g.value, 1), "passes": {}}

    sa = core.generate_scene(L.GEN_HEIGHTFIELD, 2, 1.0)
    scene = core.Scene(ctx, sa)
    s = scene.h
    scene.set_camera(core.camera_look_at(sa.eye, sa.center, sa.fov, W, H))
    img = lambda f=L.FMT_RGBA32F: core.Image(ctx, W, H, f)
    rad, acc, gp, gn, di = img(), img(), img(), img(), img()
    res = [img(), img()]
    uv = [img(L.FMT_R32U), img(L.FMT_R32U)]
    rad16, gb16, t0, t1 = (img(L.FMT_RGBA16F) for _ in range(4))
    f0, f1 = img(), img()
    ldr = img(L.FMT_RGBA8)

    p = L.PTParams(5, 0, 3, L.GUIDE_NONE, 0.0, 0, 0)
    L.call("rrh_pt_render", c, s, C.byref(p), rad.h)
    d = L.DIParams(0, 32, 1, 1, 5, 30)
    for k in (1, 0):  # two frames: the temporal stage then finds a filled previous reservoir
        L.call("rrh_di_initial", c, s, C.byref(d), gp.h, gn.h, res[k].h, uv[k].h)
    L.call("rrh_di_spatial_shade", c, s, C.byref(d), gp.h, gn.h, res[0].h, uv[0].h, di.h)
    ctx.synchronize()

    slot = C.c_int()
    ref = L.AtrousParams(5, L.ATROUS_REFERENCE, 0.25, 0.2, 1.0)
    cfg = L.AtrousParams(5, L.ATROUS_CONFIG, 1.0, 128.0, 4.0)
    ref_fast = L.AtrousParams(5, L.ATROUS_REFERENCE, 0.25, 0.2, 1.0, L.ATROUS_FAST_MATH)
    cfg_fast = L.AtrousParams(5, L.ATROUS_CONFIG, 1.0, 128.0, 4.0, L.ATROUS_FAST_MATH)
    runs = {
        "accumulate": lambda: L.call("rrh_accumulate", c, rad.h, acc.h, 0, 1),
        # reservoir 0 is re-resampled against reservoir 1 on every call; the timing is of the
        # stage's reads and writes, not of a converged reservoir
        "di_temporal": lambda: L.call("rrh_di_temporal", c, s, C.byref(d), gp.h, gn.h, res[0].h, uv[0].h, res[1].h,
                                      uv[1].h),
        "atrous_reference": lambda: L.call("rrh_atrous", c, C.byref(ref), di.h, gp.h, gn.h, f0.h, f1.h,
                                           C.byref(slot)),
        "atrous_prepare": lambda: L.call("rrh_atrous_prepare", c, s, di.h, gp.h, gn.h, rad16.h, gb16.h),
        "atrous_config": lambda: L.call("rrh_atrous", c, C.byref(cfg), rad16.h, gb16.h, None, t0.h, t1.h,
                                        C.byref(slot)),
        "atrous_reference_fast": lambda: L.call("rrh_atrous", c, C.byref(ref_fast), di.h, gp.h, gn.h, f0.h, f1.h,
                                                C.byref(slot)),
        "atrous_config_fast": lambda: L.call("rrh_atrous", c, C.byref(cfg_fast), rad16.h, gb16.h, None, t0.h, t1.h,
                                             C.byref(slot)),
        "tonemap": lambda: L.call("rrh_tonemap", c, di.h, C.c_float(1.0), ldr.h),
    }
    npix = W * H
    flusher = core.Image(ctx, 16384, 4096)   # 1 GiB RGBA32F, cleared between cold calls
    for name, fn in runs.items():
        ms, warm = _timed(ctx, reps, fn, flusher.clear)
        iters = 5 if name.startswith("atrous_") and name != "atrous_prepare" else 1
        nbytes = BYTES[name] * npix * iters
        out["passes"][name] = {"ms": round(ms, 4), "algorithmic_mb": round(nbytes / 1e6, 1),
                               "gbs": round(nbytes / (ms * 1e-3) / 1e9, 1),
                               "frac_of_copy": round(nbytes / (ms * 1e-3) / 1e9 / g.value, 3),
                               "frac_of_peak": round(nbytes / (ms * 1e-3) / 1e9 / 8000.0, 3),
                               "warm_ms": round(warm, 4)}
    flusher.close()

    # vMF grid EM (timed by a kernel trace: the entry point uploads its host arrays first)
    room = core.Scene(ctx, core.generate_scene(L.GEN_INDOOR, 4, 1.0))
    info = room.info()
    room.close()
    bmin, bmax = np.asarray(info["bounds_min"], np.float32), np.asarray(info["bounds_max"], np.float32)
    rng = np.random.default_rng(4)
    pos = (bmin + rng.random((em_samples, 3), dtype=np.float32) * (bmax - bmin)).astype(np.float32)
    v = rng.standard_normal((em_samples, 3)).astype(np.float32)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    w = (1.0 - rng.random(em_samples, dtype=np.float32)).astype(np.float32)
    grid = core.Grid(ctx, L.GridParams(L.GRID_VMM, 0, 0, 0, 0, 512.0, 32, 8, L.f3(bmin), L.f3(bmax), 0.9, 1e-6))
    fp = lambda a: a.ctypes.data_as(C.c_void_p)
    for _ in range(3):
        L.call("rrh_grid_em_step", c, grid.h, fp(pos), fp(v), fp(w), em_samples)
    ctx.synchronize()
    out["em"] = {"samples_per_step": em_samples, "steps": 3, "cells": 32 ** 3, "lobes": 8,
                 "algorithmic_bytes_per_sample": 28, "fp64_atomics_per_sample": 32}
    for o in (rad, acc, gp, gn, di, *res, *uv, rad16, gb16, t0, t1, f0, f1, ldr):
        o.close()
    grid.close()
    scene.close()
    return out


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--reps", type=int, default=20)
    a = ap.parse_args()
    ctx = core.Context(0)
    r = report(ctx, a.width, a.height, a.reps)
    print(json.dumps(r), flush=True)
    ctx.close()
    assert all(math.isfinite(x["gbs"]) for x in r["passes"].values())


if __name__ == "__main__":
    main()

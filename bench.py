#!/usr/bin/env python3
"""Headline benchmark: Mrays/s of the path tracer + ReSTIR DI on the 1,164,288-triangle
heightfield scene at 1920x1080 (1 spp per frame, slow camera orbit), on one MI355X.

One step = one frame: wavefront path tracer (1 spp, max_bounces 5, Russian roulette from
bounce 3) + accumulation + ReSTIR DI (32 candidates, temporal M-cap 20*M, 5 spatial
neighbours in a 30-pixel radius).  Rays = primary + extension + shadow rays of both
passes, counted on the device.  Inputs (scene, BVH) are resident in HBM before timing.

Prints one JSON line (see the contract in the repository notes).
"""
from __future__ import annotations

import argparse
import csv
import ctypes as C
import json
import math
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time
import types
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from research_amd import _lib as L  # noqa: E402
from research_amd import core  # noqa: E402

METRIC = "Mrays/s (1080p, 1M tris); equal-time rMSE learned guide vs vMF grid"
HBM_PEAK_GBS = 8000.0
NODE_FORMATS = {"binary": 0, "wide4": 1, "wide4q": 2}
NODE_BYTES = {"binary": 64, "wide4": 128, "wide4q": 64}
TRI_BYTES, RAY_BYTES, HIT_BYTES = 48, 32, 16


def orbit_camera(sa, frame: int, W: int, H: int):
    cx, cy, cz = sa.center
    ex, ey, ez = sa.eye
    a = math.radians(0.5 * frame)
    dx, dz = ex - cx, ez - cz
    eye = (cx + dx * math.cos(a) - dz * math.sin(a), ey, cz + dx * math.sin(a) + dz * math.cos(a))
    return core.camera_look_at(eye, sa.center, sa.fov, W, H, frame_index=frame)


class Frame:
    """One step: the path tracer (1 spp) plus accumulation, and the three ReSTIR DI stages.
    di_stream: the DI pass runs on a second context (its own HIP stream and its own copy of the
    scene), so that its kernels overlap the path tracer's on the GPU; otherwise after it on the
    same stream."""

    def __init__(self, ctx, scene, sa, W, H, max_bounces, di_m, di_ctx=None, band_ctxs=()):
        self.ctx, self.scene, self.sa, self.W, self.H = ctx, scene, sa, W, H
        self.max_bounces, self.di_m = max_bounces, di_m
        if di_ctx is not None:
            self.dctx = di_ctx
            self.dscene = core.Scene(di_ctx, sa)
        else:
            self.dctx, self.dscene = ctx, scene
        # band_ctxs: the path tracer's image split into 1 + len(band_ctxs) row bands, each traced
        # on its own context (its own stream and copy of the scene); band 0 on ctx
        self.bands = [(ctx, scene)] + [(c, core.Scene(c, sa)) for c in band_ctxs]
        n = len(self.bands)
        self.rows = [(H * k // n, H * (k + 1) // n) for k in range(n)]
        img = lambda c, f=L.FMT_RGBA32F: core.Image(c, W, H, f)
        self.rad, self.acc = img(ctx), img(ctx)
        self.gp, self.gn, self.out = img(self.dctx), img(self.dctx), img(self.dctx)
        self.res = [img(self.dctx), img(self.dctx)]
        self.uv = [img(self.dctx, L.FMT_R32U), img(self.dctx, L.FMT_R32U)]

    def pt_contexts(self):
        return [c for c, _ in self.bands]

    def contexts(self):
        return self.pt_contexts() + ([] if self.dctx is self.ctx else [self.dctx])

    def synchronize(self) -> None:
        for c in self.contexts():
            c.synchronize()

    def run(self, f: int, flags: int = 0) -> None:
        c = self.ctx.h
        cam = orbit_camera(self.sa, f, self.W, self.H)
        for _, sc in self.bands:
            sc.set_camera(cam)
        if self.dscene is not self.scene:
            self.dscene.set_camera(cam)
        whole = len(self.bands) == 1
        # each band is traced and accumulated (orbiting camera: reset every frame) on its own stream,
        # so its next frame's rays start only after its rows were accumulated, and no stream waits
        # for another
        for (bc, bs), (r0, r1) in zip(self.bands, self.rows):
            p = L.PTParams(self.max_bounces, f, 3, L.GUIDE_NONE, 0.0, 0, flags, 0, 0 if whole else r0,
                           0 if whole else r1)
            L.call("rrh_pt_render", bc.h, bs.h, C.byref(p), self.rad.h)
            if whole:
                L.call("rrh_accumulate", bc.h, self.rad.h, self.acc.h, 1, 0)
            else:
                L.call("rrh_accumulate_rows", bc.h, self.rad.h, self.acc.h, 1, 0, r0, r1)
        cur, prev = f & 1, 1 - (f & 1)
        d = L.DIParams(f, self.di_m, 1, 1, 5, 30)
        dc, ds = self.dctx.h, self.dscene.h
        L.call("rrh_di_initial", dc, ds, C.byref(d), self.gp.h, self.gn.h, self.res[cur].h, self.uv[cur].h)
        L.call("rrh_di_temporal", dc, ds, C.byref(d), self.gp.h, self.gn.h, self.res[cur].h, self.uv[cur].h,
               self.res[prev].h, self.uv[prev].h)
        L.call("rrh_di_spatial_shade", dc, ds, C.byref(d), self.gp.h, self.gn.h, self.res[cur].h, self.uv[cur].h,
               self.out.h)


def ray_totals(ctxs, reset=True):
    """rays traced since the last reset, summed over the frame's contexts"""
    tot = [0, 0, 0]
    for ctx in ctxs:
        rc = L.RayCounts()
        L.call("rrh_ray_totals", ctx.h, C.byref(rc), 1 if reset else 0)
        tot = [tot[0] + rc.primary, tot[1] + rc.extension, tot[2] + rc.shadow]
    return tuple(tot)


def span_union(spans) -> float:
    """total length of the union of [begin, end] intervals"""
    tot, cur0, cur1 = 0.0, None, None
    for b0, b1 in sorted(spans):
        if cur1 is None or b0 > cur1:
            if cur1 is not None:
                tot += cur1 - cur0
            cur0, cur1 = b0, b1
        else:
            cur1 = max(cur1, b1)
    if cur1 is not None:
        tot += cur1 - cur0
    return tot


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sa, cams, max_bounces: int, seconds: float, di_m: int) -> dict:
    """The oracle on the host cores this process may run on, for the bench's frame: ReSTIR DI's
    three stages over a whole frame (frame 1 of the orbit, its history from frame 0), and the path
    tracer on a bounded random pixel sample of the same frame, extrapolated to all pixels.  Rays
    are the oracle's own counts (primary, extension, shadow of both passes, as the GPU counts)."""
    from oracle import oracle as O

    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    osc = O.Scene.from_arrays(sa)
    W, H = cams[1].screen_width, cams[1].screen_height
    # ReSTIR DI, whole frame
    oc0, oc1 = O.camera_from_bytes(bytes(cams[0])), O.camera_from_bytes(bytes(cams[1]))
    d0, d1 = O.DIParams(0, di_m, 1, 1, 5, 30), O.DIParams(1, di_m, 1, 1, 5, 30)
    _, _, rp, rpuv = osc.di_initial(oc0, d0, threads=threads)
    t0 = time.perf_counter()
    gp, gn, rs, uv = osc.di_initial(oc1, d1, threads=threads)
    rs, uv = osc.di_temporal(oc1, d1, gp, gn, rs, uv, rp, rpuv, threads=threads)
    osc.di_spatial_shade(oc1, d1, gp, gn, rs, uv, threads=threads)
    t_di = time.perf_counter() - t0
    di_rays = W * H + int(O.lib().orc_di_last_shadow_rays())
    # path tracer, random pixels of the same frame until the time budget is spent
    rng = np.random.default_rng(0)
    n = 4096
    t_pt, pt_rays, px_done = 0.0, 0, 0
    while t_pt < max(seconds - t_di, 1.0):
        pix = rng.choice(W * H, size=n, replace=False).astype(np.uint32)
        t0 = time.perf_counter()
        _, (p, e, s) = osc.pt_render(oc1, max_bounces, 1, pix, threads=threads)
        t_pt += time.perf_counter() - t0
        pt_rays += p + e + s
        px_done += n
        n = int(min(W * H // 4, max(4096, n * 2)))
    scale = W * H / px_done
    frame_rays = pt_rays * scale + di_rays
    frame_s = t_pt * scale + t_di
    return {"value": round(frame_rays / frame_s / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "sample": f"oracle, frame 1 of the orbit at {W}x{H}: ReSTIR DI ({di_m} candidates, temporal, 5 "
                      f"neighbours) over the whole frame, {di_rays} rays in {t_di:.2f} s; path tracer (1 spp, "
                      f"max_bounces {max_bounces}) on {px_done} random pixels, {pt_rays} rays in {t_pt:.1f} s, "
                      f"extrapolated to all pixels; value = rays of one whole frame / its CPU time"}


def guiding_summary(ctx: core.Context, a) -> dict:
    """Second half of the metric: relative MSE of the unguided tracer, the 32^3 x 8-lobe vMF grid
    (online EM) and the learned MLP guide after the same wall time each, on the ~1M-triangle
    indoor scene (depth 12, 4 spp per frame, MIS 0.5, MLP batch 262,144, Adam lr 1e-3, one step
    per frame), against an FP64-accumulated reference of the same view."""
    from research_amd import guiding_eval as G
    # the MLP renders up to twice the others' time, so its time to the grid's relMSE is measured
    r = G.evaluate(ctx, a.guide_width, a.guide_height, a.guide_seconds, a.guide_ref_spp, mlp_factor=2.0)
    ms = r["methods"]
    t = a.guide_seconds
    # every method at the common time (the MLP's read off its longer curve), then the MLP's whole run
    at = {k: (G.at_time(v["curve"], t, 2), G.at_time(v["curve"], t, 3)) for k, v in ms.items()}
    return {"workload": f"indoor_{r['tri_count']}_tris_{a.guide_width}x{a.guide_height}_depth12_4spp_per_frame",
            "metric": "relMSE = mean (I - R)^2 / (R^2 + 0.01) after equal wall time",
            "seconds_per_method": t, "ref_spp": a.guide_ref_spp,
            "rel_mse": {k: v[0] for k, v in at.items()},
            "rel_mse_unbiased": {k: v[1] for k, v in at.items()},
            "spp": {k: v["spp"] for k, v in ms.items() if k != "mlp"},
            "mlp_run": {"seconds": ms["mlp"]["seconds"], "spp": ms["mlp"]["spp"], "rel_mse": ms["mlp"]["rel_mse"],
                        "rel_mse_unbiased": ms["mlp"]["rel_mse_unbiased"]},
            "mlp_over_vmm_rel_mse": r.get("mlp_over_vmm_rel_mse"),
            "mlp_over_vmm_rel_mse_unbiased": r.get("mlp_over_vmm_rel_mse_unbiased"),
            "mlp_time_to_vmm_rel_mse_ratio": r.get("mlp_time_to_vmm_rel_mse_ratio"),
            "mlp_time_to_vmm_rel_mse_unbiased_ratio": r.get("mlp_time_to_vmm_rel_mse_unbiased_ratio"),
            "configured_size": "1920x1080 with a 65,536-spp reference: profiles/guiding_equal_time_indoor_1080p.md",
            "vmm_over_unguided_rel_mse": r.get("vmm_over_unguided_rel_mse"),
            "mlp_loss": [ms["mlp"].get("loss_first"), ms["mlp"].get("loss_last")]}


def restir_pt_summary(ctx: core.Context, scene, sa, W: int, H: int, frames: int, warmup: int) -> dict:
    """ReSTIR PT (reconnection shift, the reference pass's defaults: 8 candidates of up to 4
    segments, temporal and spatial reuse, hash-grid guiding fitted every frame) over ReSTIR DI's
    G-buffer and direct image, on the same scene, extent and orbit.  Times the three ReSTIR PT
    stages plus the grid fit with markers on the context stream; rays counted on the device."""
    img = lambda f=L.FMT_RGBA32F: core.Image(ctx, W, H, f)
    gp, gn, direct, out = img(), img(), img(), img()
    dres, duv = [img(), img()], [img(L.FMT_R32U), img(L.FMT_R32U)]
    res = [[img() for _ in range(4)], [img() for _ in range(4)]]
    hist = [[img(), img()], [img(), img()]]
    grid = core.Grid(ctx, L.GridParams(L.GRID_HASH, 1 << 20, 0.10, 0.02, 8, 512.0, 0, 0, L.f3((0, 0, 0)),
                                       L.f3((1, 1, 1)), 0.0, 0.0))
    for r in dres + duv:
        r.clear()
    L.call("rrh_rpt_clear", ctx.h, core.planes(res[1]), hist[1][0].h, hist[1][1].h)
    c, s = ctx.h, scene.h
    stage_ms = np.zeros(4)
    ext = shadow = 0
    prev_vp = None
    for f in range(warmup + frames):
        cam = orbit_camera(sa, f, W, H)
        vp = list(cam.prev_view_proj)
        if prev_vp is not None:
            cam.prev_view_proj[:] = prev_vp
        prev_vp = vp
        scene.set_camera(cam)
        cur, prev = f & 1, 1 - (f & 1)
        d = L.DIParams(f, 16, 1, 1, 4, 15)
        L.call("rrh_di_initial", c, s, C.byref(d), gp.h, gn.h, dres[cur].h, duv[cur].h)
        L.call("rrh_di_temporal", c, s, C.byref(d), gp.h, gn.h, dres[cur].h, duv[cur].h, dres[prev].h, duv[prev].h)
        L.call("rrh_di_spatial_shade", c, s, C.byref(d), gp.h, gn.h, dres[cur].h, duv[cur].h, direct.h)
        p = L.RPTParams(f, 8, 4, 1, 1, 4, 12, 1, 1.0, 1e-3, 1, 0.5)
        ctx.marker(0)
        L.call("rrh_rpt_initial", c, s, C.byref(p), grid.h, gp.h, gn.h, core.planes(res[cur]), hist[cur][0].h,
               hist[cur][1].h)
        ctx.marker(1)
        L.call("rrh_rpt_temporal", c, s, C.byref(p), gp.h, gn.h, core.planes(res[cur]), core.planes(res[prev]),
               hist[prev][0].h, hist[prev][1].h)
        ctx.marker(2)
        L.call("rrh_rpt_spatial_shade", c, s, C.byref(p), grid.h, gp.h, gn.h, direct.h, core.planes(res[cur]), out.h)
        ctx.marker(3)
        L.call("rrh_grid_fit", c, grid.h)
        ctx.marker(4)
        ctx.synchronize()
        if f >= warmup:
            stage_ms += [ctx.elapsed_ms(k, k + 1) for k in range(4)]
            rc = L.RayCounts()
            L.call("rrh_rpt_ray_counts", c, C.byref(rc))
            ext += rc.extension
            shadow += rc.shadow
    ms = stage_ms / frames
    total = float(ms.sum())
    rays = (ext + shadow) / frames
    return {"workload": f"heightfield_{scene.info()['tri_count']}_tris_{W}x{H}_restir_pt_M8_4seg",
            "ms_per_frame": round(total, 3),
            "stage_ms": {"initial": round(ms[0], 3), "temporal": round(ms[1], 3), "spatial_shade": round(ms[2], 3),
                         "grid_fit": round(ms[3], 3)},
            "rays_per_frame": {"extension": ext // frames, "shadow_and_visibility": shadow // frames},
            "mrays_per_s": round(rays / (total * 1e-3) / 1e6, 1) if total > 0 else None,
            "frames": frames,
            "note": "stage times by HIP markers on the context stream; ReSTIR DI (G-buffer, direct) not included"}


def restir_gi_summary(ctx: core.Context, scene, sa, W: int, H: int, frames: int, warmup: int) -> dict:
    """ReSTIR GI (one-bounce reconnection reuse, the reference pass's defaults: 8 cosine
    candidates, temporal and spatial reuse) over ReSTIR DI's G-buffer and direct image, on the
    same scene, extent and orbit; stage times by markers on the context stream."""
    img = lambda f=L.FMT_RGBA32F: core.Image(ctx, W, H, f)
    gp, gn, direct, out = img(), img(), img(), img()
    dres, duv = [img(), img()], [img(L.FMT_R32U), img(L.FMT_R32U)]
    res = [[img() for _ in range(3)], [img() for _ in range(3)]]
    hist = [[img(), img()], [img(), img()]]
    for r in dres + duv:
        r.clear()
    L.call("rrh_rgi_clear", ctx.h, core.planes(res[1]), hist[1][0].h, hist[1][1].h)
    c, s = ctx.h, scene.h
    stage_ms = np.zeros(3)
    ext = shadow = 0
    prev_vp = None
    for f in range(warmup + frames):
        cam = orbit_camera(sa, f, W, H)
        vp = list(cam.prev_view_proj)
        if prev_vp is not None:
            cam.prev_view_proj[:] = prev_vp
        prev_vp = vp
        scene.set_camera(cam)
        cur, prev = f & 1, 1 - (f & 1)
        d = L.DIParams(f, 16, 1, 1, 4, 15)
        L.call("rrh_di_initial", c, s, C.byref(d), gp.h, gn.h, dres[cur].h, duv[cur].h)
        L.call("rrh_di_temporal", c, s, C.byref(d), gp.h, gn.h, dres[cur].h, duv[cur].h, dres[prev].h, duv[prev].h)
        L.call("rrh_di_spatial_shade", c, s, C.byref(d), gp.h, gn.h, dres[cur].h, duv[cur].h, direct.h)
        p = L.RGIParams(f, 8, 1, 1, 4, 12, 1, 1.0, 1e-3)
        ctx.marker(0)
        L.call("rrh_rgi_initial", c, s, C.byref(p), gp.h, gn.h, core.planes(res[cur]), hist[cur][0].h, hist[cur][1].h)
        ctx.marker(1)
        L.call("rrh_rgi_temporal", c, s, C.byref(p), gp.h, gn.h, core.planes(res[cur]), core.planes(res[prev]),
               hist[prev][0].h, hist[prev][1].h)
        ctx.marker(2)
        L.call("rrh_rgi_spatial_shade", c, s, C.byref(p), gp.h, gn.h, direct.h, core.planes(res[cur]), out.h)
        ctx.marker(3)
        ctx.synchronize()
        if f >= warmup:
            stage_ms += [ctx.elapsed_ms(k, k + 1) for k in range(3)]
            rc = L.RayCounts()
            L.call("rrh_rgi_ray_counts", c, C.byref(rc))
            ext += rc.extension
            shadow += rc.shadow
    ms = stage_ms / frames
    total = float(ms.sum())
    rays = (ext + shadow) / frames
    return {"workload": f"heightfield_{scene.info()['tri_count']}_tris_{W}x{H}_restir_gi_M8",
            "ms_per_frame": round(total, 3),
            "stage_ms": {"initial": round(ms[0], 3), "temporal": round(ms[1], 3), "spatial_shade": round(ms[2], 3)},
            "rays_per_frame": {"extension": ext // frames, "shadow_and_visibility": shadow // frames},
            "mrays_per_s": round(rays / (total * 1e-3) / 1e6, 1) if total > 0 else None,
            "frames": frames,
            "note": "stage times by HIP markers on the context stream; ReSTIR DI (G-buffer, direct) not included"}


# the options a traffic probe must share with the timed run (same kernels, same tree)
PROBE_OPTS = ["stack", "top_nodes", "trace_pairs", "di_stream", "pt_bands", "xcd_remap", "persistent", "refill", "suspend", "rpt_steps", "max_leaf", "trav_cost", "speculative", "bvh_builder",
              "ploc_radius", "ray_sort", "bvh_nodes", "width", "height", "max_bounces"]
COPY_BYTES = 1 << 30


def configure(ctx, a) -> None:
    L.call("rrh_ctx_set_option", ctx.h, L.OPT_TRAVERSAL_STACK,
           {"private": L.STACK_PRIVATE, "lds32": L.STACK_LDS32, "hybrid16": L.STACK_HYBRID16,
            "hybrid32": L.STACK_HYBRID32, "none": L.STACK_NONE}[a.stack])
    L.call("rrh_ctx_set_option", ctx.h, L.OPT_TRAV_TOP_NODES, a.top_nodes)
    L.call("rrh_ctx_set_option", ctx.h, L.OPT_TRACE_PAIRS, a.trace_pairs)
    L.call("rrh_ctx_set_option", ctx.h, L.OPT_XCD_REMAP, a.xcd_remap)
    L.call("rrh_ctx_set_option", ctx.h, L.OPT_BVH_NODES, NODE_FORMATS[a.bvh_nodes])
    L.call("rrh_ctx_set_option", ctx.h, L.OPT_PERSISTENT, a.persistent)
    L.call("rrh_ctx_set_option", ctx.h, L.OPT_REFILL, a.refill)
    L.call("rrh_ctx_set_option", ctx.h, L.OPT_SUSPEND_VISITS, a.suspend)
    L.call("rrh_ctx_set_option", ctx.h, L.OPT_RPT_STEP_VISITS, a.rpt_steps)
    L.call("rrh_ctx_set_option", ctx.h, L.OPT_SPECULATIVE, a.speculative)
    L.call("rrh_ctx_set_option", ctx.h, L.OPT_BVH_MAX_LEAF, a.max_leaf)
    L.call("rrh_ctx_set_option", ctx.h, L.OPT_BVH_TRAV_COST, int(round(a.trav_cost * 100)))
    L.call("rrh_ctx_set_option", ctx.h, L.OPT_BVH_BUILDER,
           L.BVH_BUILD_DEVICE if a.bvh_builder == "device" else L.BVH_BUILD_HOST)
    L.call("rrh_ctx_set_option", ctx.h, L.OPT_PLOC_RADIUS, a.ploc_radius)
    L.call("rrh_ctx_set_option", ctx.h, L.OPT_RAY_SORT, a.ray_sort)


def di_context(a):
    """the ReSTIR DI pass's own context (its own stream) when --di-stream is set"""
    if not a.di_stream:
        return None
    dctx = core.Context(0)
    configure(dctx, a)
    return dctx


def band_contexts(a):
    """contexts for the path tracer's row bands after the first (--pt-bands)"""
    out = []
    for _ in range(max(a.pt_bands, 1) - 1):
        bc = core.Context(0)
        configure(bc, a)
        out.append(bc)
    return out


def traffic_probe(a) -> int:
    """Run under rocprofv3 --pmc by measure_traffic: the copy kernel of known traffic (1 GiB in,
    1 GiB out per dispatch), then four frames of the bench's workload."""
    ctx = core.Context(0)
    configure(ctx, a)
    g = C.c_float()
    L.call("rrh_stream_copy", ctx.h, COPY_BYTES, 2, C.byref(g))
    sa = core.generate_scene(L.GEN_HEIGHTFIELD, seed=2, scale=1.0)
    scene = core.Scene(ctx, sa)
    fr = Frame(ctx, scene, sa, a.width, a.height, a.max_bounces, 32, di_context(a), band_contexts(a))
    for f in range(4):
        fr.run(f)
    fr.synchronize()
    print("[probe] done", flush=True)
    return 0


def _pmc_values(directory: str) -> dict:
    """{kernel name: [counter value per dispatch]} from rocprofv3's counter_collection csv."""
    vals: dict = {}
    for f in Path(directory).rglob("*counter_collection.csv"):
        with open(f, newline="") as fh:
            for row in csv.DictReader(fh):
                vals.setdefault(row["Kernel_Name"], []).append(float(row["Counter_Value"]))
    return vals


def _mean_of(vals: dict, pattern: str):
    xs = [v for k, vs in vals.items() if re.search(pattern, k) for v in vs]
    return sum(xs) / len(xs) if xs else None


def measure_traffic(a) -> dict:
    """HBM bytes per dispatch of this run's kernels: two rocprofv3 --pmc passes over the traffic
    probe (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch), each scaled by what the same counter
    reads for the copy kernel in the same pass."""
    exe = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(exe):
        return {"note": "rocprofv3 not found: traffic not measured"}
    extra = []
    for k in PROBE_OPTS:
        extra += ["--" + k.replace("_", "-"), str(getattr(a, k))]
    tmp = tempfile.mkdtemp(prefix="rrh_traffic_", dir="/tmp")
    per = {}
    try:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            d = os.path.join(tmp, counter)
            cmd = [exe, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "probe", "--",
                   sys.executable, str(ROOT / "bench.py"), "--traffic-probe"] + extra
            print(f"[bench] traffic pass {counter} ...", file=sys.stderr, flush=True)
            try:
                r = subprocess.run(cmd, cwd="/tmp", env={**os.environ, "TMPDIR": "/tmp"}, capture_output=True,
                                   text=True, timeout=240)
            except subprocess.TimeoutExpired:
                return {"note": f"rocprofv3 {counter} pass timed out: traffic not measured"}
            vals = _pmc_values(d)
            if "[probe] done" not in r.stdout or not vals:
                return {"note": f"rocprofv3 {counter} pass failed (rc {r.returncode}): {r.stderr.strip()[-300:]}"}
            per[counter] = vals
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    kib = COPY_BYTES / 1024.0
    cf, cw = _mean_of(per["FETCH_SIZE"], r"k_stream_copy"), _mean_of(per["WRITE_SIZE"], r"k_stream_copy")
    if not cf or not cw:
        return {"note": "copy kernel missing from the counter passes: traffic not measured"}
    fs, ws = kib / cf, kib / cw
    by_kernel = {}
    for name in sorted(set(per["FETCH_SIZE"]) & set(per["WRITE_SIZE"])):
        short = re.sub(r"\(anonymous namespace\)::|rrh::|^void\s+", "", name).split("(")[0]
        f_, w_ = _mean_of(per["FETCH_SIZE"], re.escape(name)), _mean_of(per["WRITE_SIZE"], re.escape(name))
        by_kernel[short] = round((f_ * fs + w_ * ws) * 1024.0 / 1e6, 1)
    # per launch of the roofline's kernels: every traversal launch of the path tracer
    group = r"k_intersect<false|k_trace_pair<|k_shadow<"
    fi, wi = _mean_of(per["FETCH_SIZE"], group), _mean_of(per["WRITE_SIZE"], group)
    return {"bytes": int((fi * fs + wi * ws) * 1024.0) if fi is not None and wi is not None else None,
            "fetch_scale": round(fs, 4), "write_scale": round(ws, 4), "mb_per_dispatch": by_kernel,
            "source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes of this run over 4 frames (KiB per "
                      "dispatch), scaled by the same counters for a 1 GiB copy kernel"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--max-bounces", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--stack", choices=["private", "lds32", "hybrid16", "hybrid32", "none"], default="lds32",
                    help="where the traversal stack lives (context option; none: stackless)")
    ap.add_argument("--pt-bands", type=int, default=3,
                    help="the path tracer's image in this many row bands, each on its own context and stream "
                         "(with the DI pass's, 3 bands fill the 4 hardware queues a process gets)")
    ap.add_argument("--di-stream", type=int, default=1,
                    help="1: the ReSTIR DI pass on a second context (its own stream), overlapping the path tracer")
    ap.add_argument("--trace-pairs", type=int, default=1,
                    help="1: extension rays of bounce b traced in one launch with the shadow rays of b - 1")
    ap.add_argument("--top-nodes", type=int, default=0,
                    help="top-of-tree quantized nodes each workgroup reads from LDS (0..512)")
    ap.add_argument("--xcd-remap", type=int, default=0, help="1: runs of 16 neighbouring workgroups per XCD")
    ap.add_argument("--persistent", type=int, default=0,
                    help="1: persistent queue traversal with lane refill; 2..64: at most that many waves per CU")
    ap.add_argument("--refill", type=int, default=16, help="persistent waves take new rays at this many idle lanes")
    ap.add_argument("--rpt-steps", type=int, default=0,
                    help="ReSTIR PT initial stage: node visits per traversal pass (0: whole traversals)")
    ap.add_argument("--suspend", type=int, default=L.SUSPEND_VISITS_DEFAULT,
                    help="node visits after which a ray is suspended and resumed by a second launch (0: off)")
    ap.add_argument("--max-leaf", type=int, default=8, help="largest BVH leaf (1..8)")
    ap.add_argument("--trav-cost", type=float, default=1.0, help="SAH traversal cost (triangle tests)")
    ap.add_argument("--speculative", type=int, default=1, help="1: speculative (leaf-parking) binary traversal")
    ap.add_argument("--bvh-builder", choices=["host", "device"], default="host",
                    help="BVH builder: binned SAH on the host, or PLOC on the GPU (same node formats)")
    ap.add_argument("--ploc-radius", type=int, default=16, help="device builder's neighbour window (1..32)")
    ap.add_argument("--ray-sort", type=int, default=0,
                    help="extension rays traced in key order (octant + origin Morton): key bits, 0 = off")
    ap.add_argument("--bvh-nodes", choices=list(NODE_FORMATS), default="wide4q",
                    help="BVH node format traversed (binary 64 B, 4-wide 128 B, quantized 4-wide 64 B)")
    ap.add_argument("--guide-seconds", type=float, default=4.0,
                    help="equal-time guiding comparison: wall seconds per method (0: skip)")
    ap.add_argument("--guide-ref-spp", type=int, default=8192, help="spp of the FP64 reference image")
    ap.add_argument("--guide-width", type=int, default=320)
    ap.add_argument("--guide-height", type=int, default=180)
    ap.add_argument("--restir-pt-frames", type=int, default=6,
                    help="ReSTIR PT and ReSTIR GI frames timed on the same scene and extent (0: skip)")
    ap.add_argument("--no-pass-report", action="store_true",
                    help="skip the streaming-pass report (accumulation, ReSTIR DI temporal, A-trous, tonemap)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the 1 GiB copy kernel (for hardware-counter calibration) and exit")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 counter passes that measure HBM bytes per dispatch")
    ap.add_argument("--traffic-probe", action="store_true", help=argparse.SUPPRESS)
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {a.gpus}); the multi-GPU split is "
              "described in the design notes, not implemented", file=sys.stderr)
        return 2

    if a.traffic_probe:
        return traffic_probe(a)
    ctx = core.Context(0)
    configure(ctx, a)
    if a.copy_calibration:
        g = C.c_float()
        L.call("rrh_stream_copy", ctx.h, 1 << 30, 5, C.byref(g))
        print(json.dumps({"copy_bytes_per_dispatch": 1 << 30, "copy_gbps": round(g.value, 1)}))
        return 0
    hbm = C.c_float()
    L.call("rrh_stream_copy", ctx.h, 1 << 30, 10, C.byref(hbm))
    print(f"[bench] measured copy bandwidth {hbm.value:.0f} GB/s", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    sa = core.generate_scene(L.GEN_HEIGHTFIELD, seed=2, scale=1.0)
    scene = core.Scene(ctx, sa)
    info = scene.info()
    print(f"[bench] scene: {info['tri_count']} tris, {info['node_count']} BVH nodes, depth {info['bvh_depth']}, "
          f"{info['node4_count']} 4-wide nodes (stack bound {info['stack4']}), {info['light_count']} lights, "
          f"{a.bvh_builder} BVH build+upload {info['build_ms']:.1f} ms, scene total {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)
    W, H = a.width, a.height
    fr = Frame(ctx, scene, sa, W, H, a.max_bounces, 32, di_context(a), band_contexts(a))
    for f in range(a.warmup):
        fr.run(f)
    fr.synchronize()
    # instrumented traversal on one extra frame: algorithmic bytes per closest-hit ray
    ptc = fr.pt_contexts()

    def summed(fn, k, reset=1):
        """a counter query summed over the path tracer's contexts"""
        tot = [0] * k
        for c in ptc:
            v = [C.c_uint64() for _ in range(k)]
            L.call(fn, c.h, *[C.byref(x) for x in v], reset)
            tot = [t + x.value for t, x in zip(tot, v)]
        return [types.SimpleNamespace(value=t) for t in tot]

    summed("rrh_fetch_counts", 3)
    summed("rrh_fetch_counts_shadow", 3)
    summed("rrh_fetch_wave_max", 4)
    fr.run(a.warmup, L.PT_COUNT_FETCHES)
    fr.synchronize()
    nn, nt, nr = summed("rrh_fetch_counts", 3)
    wn, wt, wsn, wst = summed("rrh_fetch_wave_max", 4)
    bytes_per_ray = (nn.value * NODE_BYTES[a.bvh_nodes] + nt.value * TRI_BYTES) / max(nr.value, 1) + RAY_BYTES + HIT_BYTES
    sn, st, sr = summed("rrh_fetch_counts_shadow", 3)
    # a shadow ray reads its origin and direction records; its contribution record and pixel are
    # touched only when it is visible, and are left out (a lower bound)
    bytes_per_shadow = ((sn.value * NODE_BYTES[a.bvh_nodes] + st.value * TRI_BYTES) / sr.value + RAY_BYTES
                        if sr.value else None)
    # the lane utilisation the spread of ray lengths alone allows: a wave lasts as long as its
    # longest ray, so sum over rays / (64 x the longest per wave), in node visits
    lane_bound = {"closest": round(nn.value / wn.value, 3) if wn.value else None,
                  "shadow": round(sn.value / wsn.value, 3) if wsn.value else None}
    print(f"[bench] traversal: {nn.value / max(nr.value, 1):.1f} nodes + {nt.value / max(nr.value, 1):.1f} tris "
          f"per ray -> {bytes_per_ray:.0f} algorithmic B/ray; shadow rays "
          f"{sn.value / max(sr.value, 1):.1f} nodes + {st.value / max(sr.value, 1):.1f} tris; lane bound from "
          f"ray lengths {lane_bound}", file=sys.stderr, flush=True)
    fr.synchronize()
    ray_totals(fr.contexts(), reset=True)
    summed("rrh_pt_traversal_rays", 2)
    for c in ptc:
        L.call("rrh_profile_enable", c.h, 1)
        L.call("rrh_profile_traversal", c.h, C.byref(C.c_float()), C.byref(C.c_uint32()), 1)
    fr.synchronize()
    L.call("rrh_profile_mark", ptc[0].h)  # time origin of the launch spans (every stream is idle here)
    t0 = time.perf_counter()
    for k in range(a.steps):
        fr.run(a.warmup + 1 + k)
    fr.synchronize()
    dt = time.perf_counter() - t0
    prim, ext, shadow = ray_totals(fr.contexts(), reset=True)
    trav_ms, launches = types.SimpleNamespace(value=0.0), types.SimpleNamespace(value=0)
    spans = []
    for c in ptc:
        L.call("rrh_profile_enable", c.h, 0)
        cap = 1 << 14
        b0, b1, ns = (C.c_float * cap)(), (C.c_float * cap)(), C.c_uint32()
        L.call("rrh_profile_spans", c.h, ptc[0].h, b0, b1, cap, C.byref(ns))
        spans += [(b0[i], b1[i]) for i in range(ns.value)]
        ms, nl = C.c_float(), C.c_uint32()
        L.call("rrh_profile_traversal", c.h, C.byref(ms), C.byref(nl), 1)
        trav_ms.value += ms.value
        launches.value += nl.value
    # the path tracer's own closest-hit and shadow rays (the ReSTIR DI pass traces its primary
    # and shadow rays inside its own kernels, which the traversal events do not cover)
    pt_closest, pt_shadow = summed("rrh_pt_traversal_rays", 2)
    rays = prim + ext + shadow
    value = rays / dt / 1e6
    avg_launch_ms = trav_ms.value / max(launches.value, 1)
    # every timed traversal launch: closest hit (bounce 0), paired closest hit + shadow, shadow
    # (the last bounce); algorithmic bytes of all their rays over all their time
    trav_bytes = bytes_per_ray * pt_closest.value + (bytes_per_shadow or 0.0) * pt_shadow.value
    bytes_per_launch = trav_bytes / max(launches.value, 1)
    achieved = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9 if avg_launch_ms > 0 else 0.0
    # the row bands' launches overlap on their own streams: the traversal's rate on the chip is all
    # their algorithmic bytes over the time at least one of them runs (the union of the spans)
    union_ms = span_union(spans)
    achieved_chip = trav_bytes / (union_ms * 1e-3) / 1e9 if union_ms > 0 else 0.0
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(dt / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene, seed 2)",
        "config": {"workload": "heightfield_1M_tris_1080p_pt_1spp_restir_di", "tri_count": info["tri_count"],
                   "width": W, "height": H, "spp_per_frame": 1, "max_bounces": a.max_bounces,
                   "di_candidates": 32, "di_temporal_mcap": "20*M", "di_spatial": "5 neighbours, r=30",
                   "camera": "orbit 0.5 deg/frame", "parallelism": "single GPU",
                   "traversal_stack": a.stack, "top_nodes_in_lds": a.top_nodes, "trace_pairs": a.trace_pairs, "di_pass_stream": "own context, overlapping the path tracer" if a.di_stream else "after the path tracer, same stream", "pt_row_bands": a.pt_bands, "xcd_remap": a.xcd_remap, "bvh_nodes": a.bvh_nodes, "persistent": a.persistent, "refill": a.refill, "suspend_visits": a.suspend, "restir_pt_step_visits": a.rpt_steps, "speculative": a.speculative,
                   "bvh_max_leaf": a.max_leaf, "bvh_trav_cost": a.trav_cost, "bvh_builder": a.bvh_builder,
                   "bvh_build_ms": round(info["build_ms"], 1), "ray_sort_bits": a.ray_sort,
                   "rays_per_step": {"primary": prim // a.steps, "extension": ext // a.steps,
                                     "shadow": shadow // a.steps}},
        "roofline": {"bound": "hbm",
                     "kernel": ("path tracer BVH traversal launches: k_intersect (bounce 0), k_trace_pair (extension "
                                "rays of bounce b with the shadow rays of b-1), k_shadow (last bounce)"
                                if a.trace_pairs else "path tracer BVH traversal launches: k_intersect, k_shadow"),
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_ray": round(bytes_per_ray, 1),
                     "bytes_per_shadow_ray": round(bytes_per_shadow, 1) if bytes_per_shadow else None,
                     "rays_per_launch": {"closest": pt_closest.value // max(launches.value, 1),
                                         "shadow": pt_shadow.value // max(launches.value, 1)},
                     "avg_launch_ms": round(avg_launch_ms, 4),
                     "launches": launches.value, "measured_copy_gbs": round(hbm.value, 1),
                     "lane_bound_from_ray_lengths": lane_bound,
                     "concurrent": {"launch_union_ms": round(union_ms, 3), "achieved": round(achieved_chip, 1),
                                    "frac": round(achieved_chip / HBM_PEAK_GBS, 4),
                                    "frame_share": round(union_ms / (dt * 1e3), 3),
                                    "note": "achieved / frac above: per launch (bytes of one launch over its own "
                                            "duration); with the image in row bands on their own streams the "
                                            "launches share the chip, so here: the algorithmic bytes of every timed "
                                            "traversal launch over the time at least one of them runs; frame_share: "
                                            "that time over the timed frames' wall time"}},
        "cpu_baseline": None,
    }
    if not a.no_traffic:
        t = measure_traffic(a)
        out["roofline"]["traffic"] = t.get("bytes")
        out["roofline"]["traffic_detail"] = t
    if a.restir_pt_frames > 0:
        print("[bench] ReSTIR PT frames ...", file=sys.stderr, flush=True)
        out["restir_pt"] = restir_pt_summary(ctx, scene, sa, W, H, a.restir_pt_frames, 2)
        print(f"[bench] ReSTIR PT {out['restir_pt']['ms_per_frame']} ms/frame, "
              f"{out['restir_pt']['mrays_per_s']} Mrays/s", file=sys.stderr, flush=True)
        out["restir_gi"] = restir_gi_summary(ctx, scene, sa, W, H, a.restir_pt_frames, 2)
        print(f"[bench] ReSTIR GI {out['restir_gi']['ms_per_frame']} ms/frame, "
              f"{out['restir_gi']['mrays_per_s']} Mrays/s", file=sys.stderr, flush=True)
    if not a.no_pass_report:
        print("[bench] streaming passes (A-trous, accumulation, ReSTIR DI temporal, tonemap) ...", file=sys.stderr,
              flush=True)
        from research_amd import kernel_report
        r = kernel_report.report(ctx, W, H, reps=10)
        out["streaming_passes"] = {
            "workload": f"heightfield_1164288_tris_{W}x{H}", "copy_gbs": r["copy_gbs"], "passes": r["passes"],
            "note": "ms per call by HIP markers around each of 10 calls, each after 1 GiB written elsewhere "
                    "(inputs start in HBM); warm_ms: 10 back-to-back calls (cache resident); algorithmic bytes "
                    "per pixel as in research_amd/kernel_report.py; A-trous passes are 5-iteration frames"}
    if a.guide_seconds > 0:
        print("[bench] equal-time guiding comparison (indoor scene) ...", file=sys.stderr, flush=True)
        out["guiding"] = guiding_summary(ctx, a)
    if not a.no_cpu_baseline:
        print("[bench] cpu baseline (oracle) ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(sa, [orbit_camera(sa, f, W, H) for f in (0, 1)], a.max_bounces,
                                           a.cpu_seconds, 32)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

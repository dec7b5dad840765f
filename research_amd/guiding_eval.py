"""Equal-time relative MSE of guided path tracing on the indoor scene (a closed
10 x 10 x 4 m room lit through one 0.5 x 0.5 m window): the unguided tracer, the
8-lobe vMF grid fitted by online EM, and the learned MLP guide, each rendering and
training for the same wall-clock budget, scored against an FP64-accumulated
high-spp reference of the same view.

relMSE = mean over pixels of mean_c (I_c - R_c)^2 / (R_c^2 + 0.01)   (eps as in
Rousselle et al. 2011 and Mueller et al. 2017).

Run:  python -m research_amd.guiding_eval --width 480 --height 270 --ref-spp 65536 --seconds 20
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys
import time

import numpy as np

from . import _lib as L
from . import core

EPS = 0.01


def rel_mse(img: np.ndarray, ref: np.ndarray) -> float:
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + EPS)))


class Setup:
    """Indoor scene, camera and the images shared by every method."""

    def __init__(self, ctx: core.Context, width: int, height: int, seed: int = 1, scale: float = 1.0,
                 max_bounces: int = 12):
        self.ctx, self.W, self.H, self.bounces = ctx, width, height, max_bounces
        sa = core.generate_scene(L.GEN_INDOOR, seed, scale)
        self.scene = core.Scene(ctx, sa)
        self.scene.set_camera(core.camera_look_at(sa.eye, sa.center, sa.fov, width, height))
        info = self.scene.info()
        self.bmin, self.bmax = info["bounds_min"], info["bounds_max"]
        self.tri_count = info["tri_count"]
        self.rad = core.Image(ctx, width, height)
        self.acc = core.Image(ctx, width, height, L.FMT_RGBA64F)

    def render(self, seed: int, mode: int = L.GUIDE_NONE, alpha: float = 0.0, record: int = 0, grid=None,
               mlp=None, guide_bounces: int = 0) -> None:
        p = L.PTParams(self.bounces, seed, 3, mode, alpha, record, 0, guide_bounces)
        L.call("rrh_pt_render_guided", self.ctx.h, self.scene.h, C.byref(p), grid.h if grid else None,
               mlp.h if mlp else None, self.rad.h)

    def accumulate(self, spp_before: int) -> None:
        L.call("rrh_accumulate_f64", self.ctx.h, self.rad.h, self.acc.h, 0, spp_before)

    def mean_image(self, spp: int) -> np.ndarray:
        return self.acc.download()[..., :3] / float(spp)

    def reference(self, spp: int, seed0: int = 1 << 24, progress: bool = False) -> np.ndarray:
        for k in range(spp):
            self.render(seed0 + k)
            self.accumulate(k)
            if progress and (k + 1) % 4096 == 0:
                self.ctx.synchronize()
                print(f"reference: {k + 1}/{spp} spp", file=sys.stderr, flush=True)
        return self.mean_image(spp)


def run_method(S: Setup, method: str, seconds: float, ref: np.ndarray, spp_per_frame: int = 4,
               alpha: float = 0.5, batch: int = 262144, record: int | None = None, train_steps: int = 1,
               checkpoints: int = 12, seed0: int = 0, lr: float = 1e-3, decay: float = 0.95,
               min_weight: float = 0.0, res: int = 32, guide_bounces: int = 0) -> dict:
    """Render (and train) for `seconds` of wall time; relMSE at evenly spaced checkpoints.
    The clock stops while a checkpoint image is read back and scored.  The grid keeps its EM
    statistics across frames with weight `decay` per frame (0.95: about 20 frames of memory)."""
    ctx = S.ctx
    grid = mlp = None
    if record is None:  # vertices kept per frame: ~2M for the EM fit, one batch for the MLP
        record = batch if method == "mlp" else 2_000_000
    if method == "vmm":
        prm = L.GridParams(L.GRID_VMM, 0, 0, 0, 0, 512.0, res, 8, L.f3(S.bmin), L.f3(S.bmax), decay, min_weight)
        grid = core.Grid(ctx, prm)
        mode = L.GUIDE_VMM
    elif method == "mlp":
        mlp = core.Mlp(ctx, S.bmin, S.bmax, lr=lr, seed=1)
        mode = L.GUIDE_MLP
    else:
        mode = L.GUIDE_NONE
    ctx.synchronize()
    curve, spp, frame, elapsed = [], 0, 0, 0.0
    next_cp = seconds / checkpoints
    losses = []
    t0 = time.perf_counter()
    while True:
        a = alpha if (mode != L.GUIDE_NONE and frame > 0) else 0.0
        for s in range(spp_per_frame):
            rec = record if (mode != L.GUIDE_NONE and s == 0) else 0
            S.render(seed0 + spp, mode, a, rec, grid, mlp, guide_bounces)
            S.accumulate(spp)
            spp += 1
        if method == "vmm":
            L.call("rrh_grid_em_from_vertices", ctx.h, grid.h)
        elif method == "mlp":
            for _ in range(train_steps):
                losses.append(mlp.train_from_vertices(batch))
        frame += 1
        ctx.synchronize()
        now = elapsed + (time.perf_counter() - t0)
        if now >= next_cp or now >= seconds:
            elapsed = now
            curve.append((round(now, 4), spp, rel_mse(S.mean_image(spp), ref)))
            next_cp += seconds / checkpoints
            if now >= seconds:
                break
            t0 = time.perf_counter()
    out = {"method": method, "seconds": round(elapsed, 3), "spp": spp, "frames": frame, "rel_mse": curve[-1][2],
           "curve": curve}
    if losses:
        out["loss_first"], out["loss_last"] = float(losses[0]), float(np.mean(losses[-5:]))
    return out


def time_to_reach(curve, target: float) -> float | None:
    """First time at which a relMSE curve reaches `target` (linear in log relMSE between checkpoints)."""
    prev = None
    for t, _, e in curve:
        if e <= target:
            if prev is None:
                return t
            t0, e0 = prev
            if e0 <= e:
                return t
            f = (np.log(e0) - np.log(target)) / (np.log(e0) - np.log(e))
            return float(t0 + f * (t - t0))
        prev = (t, e)
    return None


def evaluate(ctx: core.Context, width: int = 320, height: int = 180, seconds: float = 4.0, ref_spp: int = 8192,
             methods=("none", "vmm", "mlp"), scale: float = 1.0, progress: bool = False, **kw) -> dict:
    S = Setup(ctx, width, height, scale=scale)
    t = time.perf_counter()
    ref = S.reference(ref_spp, progress=progress)
    ctx.synchronize()
    ref_s = time.perf_counter() - t
    res = {"scene": "indoor_10x10x4_window_sun", "tri_count": S.tri_count, "width": width, "height": height,
           "max_bounces": S.bounces, "ref_spp": ref_spp, "ref_seconds": round(ref_s, 2),
           "ref_mean": float(ref.mean()), "ref_lit_fraction": float((ref.sum(-1) > 0).mean()), "seconds": seconds,
           "methods": {}}
    for m in methods:
        res["methods"][m] = run_method(S, m, seconds, ref, **kw)
        if progress:
            r = res["methods"][m]
            print(f"{m}: relMSE {r['rel_mse']:.5f} at {r['spp']} spp in {r['seconds']} s", file=sys.stderr, flush=True)
    ms = res["methods"]
    if "vmm" in ms and "mlp" in ms:
        res["mlp_over_vmm_rel_mse"] = ms["mlp"]["rel_mse"] / ms["vmm"]["rel_mse"]
        tv = ms["vmm"]["seconds"]
        tm = time_to_reach(ms["mlp"]["curve"], ms["vmm"]["rel_mse"])
        res["mlp_time_to_vmm_rel_mse_ratio"] = None if tm is None else tm / tv
    if "none" in ms and "vmm" in ms:
        res["vmm_over_unguided_rel_mse"] = ms["vmm"]["rel_mse"] / ms["none"]["rel_mse"]
    return res


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--width", type=int, default=480)
    ap.add_argument("--height", type=int, default=270)
    ap.add_argument("--seconds", type=float, default=20.0)
    ap.add_argument("--ref-spp", type=int, default=65536)
    ap.add_argument("--methods", default="none,vmm,mlp")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--spp-per-frame", type=int, default=4)
    ap.add_argument("--batch", type=int, default=262144)
    ap.add_argument("--train-steps", type=int, default=1)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--alpha", type=float, default=0.5)
    ap.add_argument("--guide-bounces", type=int, default=0, help="guide only the first N bounces (0: all)")
    a = ap.parse_args()
    ctx = core.Context(0)
    res = evaluate(ctx, a.width, a.height, a.seconds, a.ref_spp, tuple(a.methods.split(",")), a.scale, True,
                   spp_per_frame=a.spp_per_frame, batch=a.batch, train_steps=a.train_steps, lr=a.lr, alpha=a.alpha,
                   guide_bounces=a.guide_bounces)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()

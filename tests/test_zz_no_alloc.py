"""rrh_pt_render_noalloc (no device allocation: queues and counters in a file-scope device
array) against rrh_pt_render on the same scene, camera and seed: images and ray counts bit for
bit, on two calls in a row, the second with a smaller extent, and the size limit enforced."""
import ctypes as C

import numpy as np
import pytest

from research_amd import _lib as L
from research_amd import core

pytestmark = pytest.mark.gpu


def _frame(ctx, gs, W, H, seed, depth, fn):
    img = core.Image(ctx, W, H)
    p = L.PTParams(depth, seed, 3, 0, 0.0, 0)
    print(f"no-alloc test: {fn} {W}x{H} seed {seed} depth {depth}", flush=True)
    L.call(fn, ctx.h, gs.h, C.byref(p), img.h)
    rc = L.RayCounts()
    L.call("rrh_pt_ray_counts", ctx.h, C.byref(rc))
    return img.download(), (rc.primary, rc.extension, rc.shadow)


def test_noalloc_render_equals_the_allocating_entry_point(ctx):
    cases = [  # second call smaller than the first; the last one at the 512 x 512 limit
        (L.GEN_CORNELL_RANDOM, 1, 1.0, 320, 240, 3, 8),
        (L.GEN_HEIGHTFIELD, 2, 0.1, 96, 64, 5, 6),
        (L.GEN_CORNELL_REF, 1, 1.0, 512, 512, 0, 5),
    ]
    for kind, seed, scale, W, H, frame, depth in cases:
        sa = core.generate_scene(kind, seed, scale)
        gs = core.Scene(ctx, sa)
        gs.set_camera(core.camera_look_at(sa.eye, sa.center, sa.fov, W, H))
        g, gc = _frame(ctx, gs, W, H, frame, depth, "rrh_pt_render_noalloc")
        r, rc = _frame(ctx, gs, W, H, frame, depth, "rrh_pt_render")
        assert gc == rc and gc[0] == W * H, (gc, rc)
        assert np.array_equal(g.view(np.uint32), r.view(np.uint32)), (kind, W, H)
        # and again after the allocating call, on the same array
        g2, gc2 = _frame(ctx, gs, W, H, frame, depth, "rrh_pt_render_noalloc")
        assert gc2 == rc and np.array_equal(g2.view(np.uint32), r.view(np.uint32))


def test_noalloc_render_rejects_what_it_cannot_do(ctx):
    sa = core.generate_scene(L.GEN_CORNELL_REF)
    gs = core.Scene(ctx, sa)
    gs.set_camera(core.camera_look_at(sa.eye, sa.center, sa.fov, 640, 480))
    img = core.Image(ctx, 640, 480)
    p = L.PTParams(5, 0, 3, 0, 0.0, 0)
    with pytest.raises(L.RRHError, match="above the no-alloc limit of 262144 pixels"):
        L.call("rrh_pt_render_noalloc", ctx.h, gs.h, C.byref(p), img.h)
    gs.set_camera(core.camera_look_at(sa.eye, sa.center, sa.fov, 64, 64))
    img = core.Image(ctx, 64, 64)
    for bad in (L.PTParams(5, 0, 3, L.GUIDE_VMM, 0.5, 0), L.PTParams(5, 0, 3, 0, 0.0, 1000),
                L.PTParams(5, 0, 3, 0, 0.0, 0, L.PT_COUNT_FETCHES)):
        with pytest.raises(L.RRHError, match="rrh_pt_render_noalloc"):
            L.call("rrh_pt_render_noalloc", ctx.h, gs.h, C.byref(bad), img.h)
    L.call("rrh_ctx_set_option", ctx.h, L.OPT_RAY_SORT, 16)
    try:
        with pytest.raises(L.RRHError, match="ray sorting"):
            L.call("rrh_pt_render_noalloc", ctx.h, gs.h, C.byref(p), img.h)
    finally:
        L.call("rrh_ctx_set_option", ctx.h, L.OPT_RAY_SORT, 0)

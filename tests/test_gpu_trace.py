"""GPU parity of scene flattening and ray queries against the oracle (bit-exact)."""
import numpy as np
import pytest

from oracle import oracle as O
from research_amd import _lib as L
from research_amd import core
from tests.helpers import prim_ids, primary_rays, random_rays

pytestmark = pytest.mark.gpu


def _scene(ctx, kind, seed=1, scale=1.0):
    sa = core.generate_scene(kind, seed, scale)
    return sa, core.Scene(ctx, sa), O.Scene.from_arrays(sa)


@pytest.mark.parametrize("kind,scale", [(L.GEN_CORNELL_REF, 1.0), (L.GEN_CORNELL_RANDOM, 1.0),
                                        (L.GEN_HEIGHTFIELD, 0.08), (L.GEN_INDOOR, 0.05)])
def test_flattened_triangles_match_oracle(ctx, kind, scale):
    sa, gs, os_ = _scene(ctx, kind, 3, scale)
    a, b = gs.triangles(), os_.triangles()
    assert a.shape == b.shape
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    info = gs.info()
    assert info["tri_count"] == sa.tri_count and info["bvh_depth"] < 64 and info["stack4"] <= 64


def test_trace_cornell_bit_exact(ctx):
    sa, gs, os_ = _scene(ctx, L.GEN_CORNELL_REF)
    cam = core.camera_look_at(sa.eye, sa.center, sa.fov, 64, 64)
    rays = np.concatenate([primary_rays(cam), random_rays(50000, -0.99, 0.99, 1)])
    g = core.trace_closest(ctx, gs, rays)
    o = os_.trace(rays, brute=True)
    assert np.array_equal(g.view(np.uint32), o.view(np.uint32))


@pytest.mark.parametrize("kind,scale,lo,hi", [(L.GEN_HEIGHTFIELD, 0.25, [-45, 0, -45], [45, 20, 45]),
                                              (L.GEN_INDOOR, 0.2, [0.1, 0.1, 0.1], [9.9, 3.9, 9.9])])
def test_trace_procedural_bit_exact(ctx, kind, scale, lo, hi):
    sa, gs, os_ = _scene(ctx, kind, 2, scale)
    rays = random_rays(100000, lo, hi, 4)
    g = core.trace_closest(ctx, gs, rays)
    o = os_.trace(rays)
    pg, po = prim_ids(g), prim_ids(o)
    assert (pg == po).mean() == 1.0
    hit = po != 0xFFFFFFFF
    assert np.array_equal(g[hit].view(np.uint32), o[hit].view(np.uint32))
    occ = core.trace_occluded(ctx, gs, rays)
    assert np.array_equal(occ.astype(bool), hit)


def test_trace_full_size_heightfield_primary_rays(ctx):
    # the 1,164,288-triangle scene at 1920x1080: a 1-in-97 sample of the primary rays
    sa, gs, os_ = _scene(ctx, L.GEN_HEIGHTFIELD, 2, 1.0)
    info = gs.info()
    assert info["tri_count"] == 1_164_288 and info["bvh_depth"] < 64
    cam = core.camera_look_at(sa.eye, sa.center, sa.fov, 1920, 1080)
    import ctypes as C
    l = O.lib()
    oc = O.camera_from_bytes(bytes(cam))
    pix = np.arange(0, 1920 * 1080, 97, dtype=np.int64)
    rays = np.zeros((len(pix), 8), np.float32)
    o3, d3 = (C.c_float * 3)(), (C.c_float * 3)()
    for k, pid in enumerate(pix):
        px, py = int(pid % 1920), int(pid // 1920)
        s = C.c_uint32(l.orc_make_seed(px, py, 0))
        l.orc_primary_ray(C.byref(oc), px, py, C.byref(s), o3, d3)
        rays[k] = [o3[0], o3[1], o3[2], 0.001, d3[0], d3[1], d3[2], 1e9]
    g = core.trace_closest(ctx, gs, rays)
    o = os_.trace(rays)
    match = (prim_ids(g) == prim_ids(o)).mean()
    assert match >= 0.9999
    hit = prim_ids(o) != 0xFFFFFFFF
    rel = np.abs(g[hit, 0] - o[hit, 0]) / np.abs(o[hit, 0])
    assert rel.max() <= 1e-5


@pytest.mark.parametrize("kind,seed,scale,lo,hi", [(L.GEN_HEIGHTFIELD, 2, 1.0, [-60, 0, -60], [60, 25, 60]),
                                                   (L.GEN_INDOOR, 1, 1.0, [0.1, 0.1, 0.1], [9.9, 3.9, 9.9])])
def test_node_formats_agree(ctx, kind, seed, scale, lo, hi):
    """The three node formats of the same tree (binary, also with speculative traversal; 4-wide
    fp32; 4-wide quantized), the stackless traversal and the top of the tree read from LDS give
    identical closest hits and occlusion on the full-size scenes; the 4-wide tree keeps its
    stack bound."""
    sa = core.generate_scene(kind, seed, scale)
    gs = core.Scene(ctx, sa)
    info = gs.info()
    print(kind, {k: info[k] for k in ("tri_count", "node_count", "bvh_depth", "node4_count", "stack4")})
    assert 0 < info["node4_count"] < info["node_count"] and info["stack4"] <= 64
    assert info["bvh_depth4"] <= 31     # the stackless traversal's reach
    rays = random_rays(200000, lo, hi, 9)
    res = {}
    P = L.STACK_PRIVATE
    # (node format, speculative, stack kind, top-of-tree nodes in LDS); the stackless traversal
    # and the LDS top only apply to 4-wide trees (the top to quantized nodes)
    combos = [(L.NODES_BINARY, 0, P, 0), (L.NODES_BINARY, 1, P, 0), (L.NODES_WIDE4, 0, P, 0),
              (L.NODES_WIDE4Q, 0, P, 0), (L.NODES_WIDE4Q, 1, P, 0), (L.NODES_WIDE4Q, 1, P, 85),
              (L.NODES_WIDE4Q, 1, P, 512), (L.NODES_WIDE4, 1, L.STACK_NONE, 0),
              (L.NODES_WIDE4Q, 1, L.STACK_NONE, 0), (L.NODES_WIDE4Q, 1, L.STACK_NONE, 21),
              (L.NODES_WIDE4Q, 1, L.STACK_NONE, 512)]
    try:
        for fmt, spec, stack, top in combos:
            L.call("rrh_ctx_set_option", ctx.h, L.OPT_BVH_NODES, fmt)
            L.call("rrh_ctx_set_option", ctx.h, L.OPT_SPECULATIVE, spec)
            L.call("rrh_ctx_set_option", ctx.h, L.OPT_TRAVERSAL_STACK, stack)
            L.call("rrh_ctx_set_option", ctx.h, L.OPT_TRAV_TOP_NODES, top)
            res[(fmt, spec, stack, top)] = core.trace_closest(ctx, gs, rays), core.trace_occluded(ctx, gs, rays)
    finally:  # back to the defaults
        L.call("rrh_ctx_set_option", ctx.h, L.OPT_BVH_NODES, L.NODES_WIDE4Q)
        L.call("rrh_ctx_set_option", ctx.h, L.OPT_SPECULATIVE, 1)
        L.call("rrh_ctx_set_option", ctx.h, L.OPT_TRAVERSAL_STACK, L.STACK_LDS32)
        L.call("rrh_ctx_set_option", ctx.h, L.OPT_TRAV_TOP_NODES, 0)
    g0, o0 = res[combos[0]]
    for key in combos[1:]:
        assert np.array_equal(res[key][0].view(np.uint32), g0.view(np.uint32)), key
        assert np.array_equal(res[key][1], o0), key
    assert 0.05 < (prim_ids(g0) != 0xFFFFFFFF).mean()
    assert np.array_equal(o0.astype(bool), prim_ids(g0) != 0xFFFFFFFF)

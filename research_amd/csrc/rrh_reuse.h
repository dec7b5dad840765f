// rrh_reuse.h -- pieces shared by the reservoir-reuse passes that reconnect a secondary vertex
// to the primary surface: ReSTIR PT (rrh_restir_pt.hip, restir_pt.slang) and ReSTIR GI
// (rrh_restir_gi.hip, restir_gi.slang).  The two shaders define these helpers identically:
// light sampling, reconnection Jacobian, visibility, the G-buffer and history surfaces,
// reprojection into the previous frame and the spatial neighbour pattern.
#pragma once
#include <algorithm>
#include <string>

#include "rrh_internal.h"

namespace rrh {
namespace reuse {

// Kernel template flag: rays cull back faces (RRH_OPT_RESTIR_CULL).  The shaders ask for
// RAY_FLAG_CULL_BACK_FACING_TRIANGLES (restir_pt.slang:104,111, restir_gi.slang:75,82), but the
// reference's TLAS instances carry TRIANGLE_FACING_CULL_DISABLE (AccelStructure.cpp:90), so by
// default every ray is two-sided; the flag rides in the kernels' stack-kind template argument.
constexpr int kCull = 16;

// trace_ray as an occlusion query (CULL: only triangles facing the ray occlude)
template <bool CULL, class Stk>
RRH_D bool occluded(const DevScene& S, f3 o, f3 d, float tmin, float tmax, Stk& st) {
    if (S.node_fmt == RRH_NODES_WIDE4) return trace_occluded4_s<false, CULL>(S, o, d, tmin, tmax, st);
    return trace_occluded4_s<true, CULL>(S, o, d, tmin, tmax, st);
}

// sample_lights (restir_pt.slang:171-222, restir_gi.slang:143-193) up to its shadow ray: c = the
// contribution if the light is visible, and the shadow ray (origin, direction, tmax; tmin =
// bias).  false: no ray is needed and the contribution is zero.  The shadow test only decides
// whether c counts, so c is formed before the ray is traced.  Triangle lights (new; the
// reference samples point and directional lights only) are sampled by area, without MIS: these
// passes never add emission at a BSDF hit beyond the reconnection vertex.
RRH_D bool light_sample(const DevScene& S, f3 hit, f3 n, f3 wo, const Mat& m, float bias, uint32_t& seed, f3& c,
                        f3& ro, f3& rd, float& tmax) {
    c = mk(0.0f, 0.0f, 0.0f);
    const uint32_t N = S.nl + S.ntl;
    if (N == 0) return false;
    const uint32_t li = pick_index(rnd(seed), N);
    if (li < S.nl) {
        const rrh_light& L = S.lights[li];
        f3 ldir;
        float dist, pl;
        if (L.type == 0) {
            const f3 to = sub(ld3(L.position), hit);
            dist = length(to);
            ldir = divs(to, fmaxf(dist, 1e-6f));
            pl = dist * dist;
        } else {
            ldir = normalize(neg(ld3(L.direction)));
            dist = 1e9f;
            pl = 1.0f;
        }
        if (dot(ldir, n) <= 0.0f) return false;
        const f3 f = mat_eval(m, n, wo, ldir);
        const float ndl = fmaxf(dot(n, ldir), 0.0f);
        const f3 irr = divs(scale(ld3(L.emission), L.intensity), (L.type == 0) ? (pl + 1e-6f) : 1.0f);
        const float pb = mat_ref_pdf(m, wo, ldir, n);
        const float mis = (L.type == 0) ? 1.0f : (pl * pl) / (pl * pl + pb * pb + 1e-10f);
        c = scale(scale(scale(mul(f, irr), ndl), mis), (float)N);
        ro = add(hit, scale(n, bias));
        rd = ldir;
        tmax = dist - 0.01f;
        return true;
    }
    const float u1 = rnd(seed);
    const float u2 = rnd(seed);
    const TriSample ts = tri_light_point(S, li - S.nl, u1, u2);
    const f3 to = sub(ts.p, hit);
    const float d2 = dot(to, to);
    const float dist = sqrtf(d2);
    const f3 l = scale(to, 1.0f / dist);
    if (dot(l, n) <= 0.0f) return false;
    const float cosl = fabsf(dot(ts.nl, l));
    if (!(cosl > 0.0f)) return false;
    const f3 f = mat_eval(m, n, wo, l);
    const float ndl = fmaxf(dot(n, l), 0.0f);
    const float pl = d2 / (cosl * ts.area) / (float)N;
    c = scale(scale(mul(f, ts.Le), ndl), 1.0f / pl);
    ro = add(hit, scale(n, bias));
    rd = l;
    tmax = dist - 0.01f;
    return true;
}

// reconnection_jacobian (restir_pt.slang:427-447, sample_reconnection_jacobian
// restir_gi.slang:461-485), with 1/sqrtf for the shaders' rsqrt
RRH_D float jacobian(f3 src, f3 dst, f3 rc, f3 rn) {
    const f3 sd = sub(rc, src), td = sub(rc, dst);
    const float s2 = dot(sd, sd), t2 = dot(td, td);
    if (s2 <= 1e-8f || t2 <= 1e-8f) return 0.0f;
    const f3 sdir = scale(sd, 1.0f / sqrtf(s2)), tdir = scale(td, 1.0f / sqrtf(t2));
    const float sc = fabsf(dot(rn, neg(sdir))), tc = fabsf(dot(rn, neg(tdir)));
    if (sc <= 0.05f || tc <= 0.05f) return 0.0f;
    return clampf((tc * s2) / (sc * t2), 0.0f, 3.0f);
}

// trace_visibility (restir_pt.slang:449-458, restir_gi.slang:553-564)
template <bool CULL, class Stk>
RRH_D bool visible(const DevScene& S, f3 o, f3 n, f3 target_pos, float bias, Stk& st) {
    const f3 delta = sub(target_pos, o);
    const float dist = length(delta);
    if (dist <= bias) return false;
    return !occluded<CULL>(S, add(o, scale(n, bias)), divs(delta, dist), bias, dist - bias, st);
}

// load_surface_from_images (restir_pt.slang:285-305, restir_gi.slang:253-276): w < 0 is a miss
RRH_D bool load_surface(const float4* pos, const float4* nrm, size_t i, f3& p, f3& n, uint32_t& mat) {
    const float4 a = pos[i];
    if (a.w < 0.0f) return false;
    const float4 b = nrm[i];
    p = xyz(a);
    n = normalize(xyz(b));
    mat = __float_as_uint(b.w);
    return true;
}

struct Frame {
    const float4* gpos;
    const float4* gnrm;
    uint32_t W, H;
};

RRH_D bool tile_pixel(uint32_t& px, uint32_t& py, uint32_t W, uint32_t H, bool remap) {
    const uint32_t tile = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y, remap);
    px = (tile % gridDim.x) * 8 + (threadIdx.x & 7u);
    py = (tile / gridDim.x) * 8 + (threadIdx.x >> 3);
    return px < W && py < H;
}

// rays of a wave summed into one counter (one atomic per wave)
RRH_D void count_rays(uint32_t* counter, uint32_t n) {
    for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off);
    if ((threadIdx.x & 63u) == 0 && n) atomicAdd(counter, n);
}

struct History {
    const float4* pos;
    const float4* nrm;
};

RRH_D bool reproject(const rrh_camera& cam, f3 p, uint32_t W, uint32_t H, int& hx, int& hy) {
    const float* M = cam.prev_view_proj;
    const float cx = mrow(M, 0, p.x, p.y, p.z, 1.0f), cy = mrow(M, 1, p.x, p.y, p.z, 1.0f),
                cw = mrow(M, 3, p.x, p.y, p.z, 1.0f);
    if (cw <= 1.0e-6f) return false;
    const float nx = cx / cw, ny = cy / cw;
    if (fabsf(nx) > 1.0f || fabsf(ny) > 1.0f) return false;
    const float fx = (nx * 0.5f + 0.5f) * (float)W - 0.5f, fy = (ny * 0.5f + 0.5f) * (float)H - 0.5f;
    hx = (int)floorf(fx + 0.5f);
    hy = (int)floorf(fy + 0.5f);
    return hx >= 0 && hy >= 0 && hx < (int)W && hy < (int)H;
}

// find_reprojected_history_surface (restir_pt.slang:355-401, restir_gi.slang:358-433): of
// the 3x3 previous-frame pixels around the reprojection, the compatible surface (same
// material, normals within 0.9, plane and position deltas within 1% and 5% of the view
// distance) with the smallest plane x 4 + position error
RRH_D bool find_history(const rrh_camera& cam, const History& Hp, uint32_t W, uint32_t H, f3 pp, f3 pn, uint32_t pmi,
                        f3 cp, size_t& hj, f3& hp) {
    int cx, cy;
    if (!reproject(cam, pp, W, H, cx, cy)) return false;
    bool found = false;
    float best = 1e30f;
    hj = 0;
    hp = mk(0.0f, 0.0f, 0.0f);
    const float vdist = fmaxf(length(sub(pp, cp)), 1.0f);
    for (int oy = -1; oy <= 1; ++oy)
        for (int ox = -1; ox <= 1; ++ox) {
            const int qx = cx + ox, qy = cy + oy;
            if (qx < 0 || qy < 0 || qx >= (int)W || qy >= (int)H) continue;
            const size_t j = (size_t)qy * W + (size_t)qx;
            f3 q, qn;
            uint32_t qm;
            if (!load_surface(Hp.pos, Hp.nrm, j, q, qn, qm)) continue;
            if (qm != pmi || dot(pn, qn) < 0.9f) continue;
            const float plane = fabsf(dot(sub(q, pp), pn));
            const float dpos = length(sub(q, pp));
            if (!(plane <= 0.01f * vdist && dpos <= 0.05f * vdist)) continue;
            const float err = plane * 4.0f + dpos;
            if (err < best) { best = err; hj = j; hp = q; found = true; }
        }
    return found;
}

RRH_D float fracf(float x) { return x - floorf(x); }

// the k-th spatial neighbour (restir_pt.slang:905-918, restir_gi.slang:907-923): a golden-ratio
// spiral of radius sqrt((k + 1/2) / n) x spatial_radius turned by the pixel's hashed phase,
// clamped to the image; false when it lands on the pixel itself
RRH_D bool neighbour(uint32_t k, uint32_t n, uint32_t radius_px, float base_phase, uint32_t px, uint32_t py,
                     uint32_t W, uint32_t H, size_t& j) {
    const float u = ((float)k + 0.5f) / (float)n;
    const float angle = fracf(base_phase + 0.61803398875f * (float)k) * K_TWO_PI;
    const float radius = sqrtf(u) * (float)radius_px;
    float sa, ca;
    sincos_rr(angle, &sa, &ca);
    int nx = (int)px + (int)(radius * ca), ny = (int)py + (int)(radius * sa);
    nx = nx < 0 ? 0 : (nx > (int)W - 1 ? (int)W - 1 : nx);
    ny = ny < 0 ? 0 : (ny > (int)H - 1 ? (int)H - 1 : ny);
    if (nx == (int)px && ny == (int)py) return false;
    j = (size_t)ny * W + (size_t)nx;
    return true;
}
RRH_D float spiral_phase(uint32_t px, uint32_t py) {
    return fracf((float)pcg_hash(px ^ (py * 65537u)) * 2.3283064365386963e-10f);
}

// is_spatial_neighbor_compatible (restir_pt.slang:403-421, restir_gi.slang:435-459)
RRH_D bool neighbour_surface(const Frame& F, size_t j, f3 pp, f3 pn, float dist_scale, f3& q) {
    const float4 q4 = F.gpos[j];
    if (q4.w < 0.0f) return false;
    q = xyz(q4);
    const f3 qn = normalize(xyz(F.gnrm[j]));
    if (dot(qn, pn) < 0.85f) return false;
    return fabsf(dot(sub(q, pp), pn)) <= 0.02f * dist_scale;
}

// ---- host side ---------------------------------------------------------------------------
inline bool fits(const rrh_image* a, uint32_t W, uint32_t H) { return a && a->w == W && a->h == H && a->fmt == RRH_FMT_RGBA32F; }

// the reuse passes' argument checks (the ray counters are cleared by each initial stage)
inline int check_frame(rrh_ctx* c, rrh_scene* s, const rrh_image* gpos, const rrh_image* gnrm, uint32_t candidates,
                       uint32_t neighbours, Frame& F, const char* who) {
    if (!(c && s && gpos && gnrm)) { set_error(std::string(who) + ": null argument"); return RRH_E_ARG; }
    if (!s->cam_set) { set_error(std::string(who) + ": camera not set (rrh_camera_set)"); return RRH_E_ARG; }
    const uint32_t W = s->cam_host.screen_width, H = s->cam_host.screen_height;
    if (!fits(gpos, W, H) || !fits(gnrm, W, H)) {
        set_error(std::string(who) + ": G-buffer position / normal must be RGBA32F images of the camera's extent");
        return RRH_E_ARG;
    }
    if (candidates > 4096 || neighbours > 64) {
        set_error(std::string(who) + ": num_initial_candidates <= 4096, spatial_num_neighbors <= 64");
        return RRH_E_RANGE;
    }
    if (s->stack4 > RRH_STACK) { set_error(std::string(who) + ": tree deeper than the traversal stack"); return RRH_E_STATE; }
    F = Frame{(const float4*)gpos->p, (const float4*)gnrm->p, W, H};
    return RRH_OK;
}

// the kernels' template argument: stack kind (these rays traverse 4-wide nodes, so the stack
// must hold what the 4-wide tree can need) plus kCull when back faces are culled
inline int kernel_kind(const rrh_ctx* c, const rrh_scene* s) {
    return stack_for(c, s->stack4) | (c->restir_cull ? kCull : 0);
}

// calls f with the kernel K<kind>::fn for a run-time kind
template <template <int> class K, class Fn>
inline int with_kernel(int kind, Fn&& f) {
    switch (kind) {
        case RRH_STACK_LDS32: return f(K<RRH_STACK_LDS32>::fn);
        case RRH_STACK_HYBRID16: return f(K<RRH_STACK_HYBRID16>::fn);
        case RRH_STACK_HYBRID32: return f(K<RRH_STACK_HYBRID32>::fn);
        case RRH_STACK_PRIVATE | kCull: return f(K<RRH_STACK_PRIVATE | kCull>::fn);
        case RRH_STACK_LDS32 | kCull: return f(K<RRH_STACK_LDS32 | kCull>::fn);
        case RRH_STACK_HYBRID16 | kCull: return f(K<RRH_STACK_HYBRID16 | kCull>::fn);
        case RRH_STACK_HYBRID32 | kCull: return f(K<RRH_STACK_HYBRID32 | kCull>::fn);
        default: return f(K<RRH_STACK_PRIVATE>::fn);
    }
}

template <template <int> class K, class... A>
void launch(int kind, dim3 g, hipStream_t st, A... a) {
    with_kernel<K>(kind, [&](auto fn) {
        hipLaunchKernelGGL(fn, g, dim3(64), 0, st, a...);
        return 0;
    });
}

// persistent waves for a kernel: as many as the CUs hold at once, never more than `tiles`
template <template <int> class K>
inline int persistent_waves(const rrh_ctx* c, int kind, uint32_t tiles, uint32_t& waves) {
    return with_kernel<K>(kind, [&](auto fn) {
        int per_cu = 0;
        RRH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 64, 0));
        waves = std::min<uint32_t>(tiles, (uint32_t)c->cu_count * (uint32_t)std::max(per_cu, 1));
        return RRH_OK;
    });
}

}  // namespace reuse
}  // namespace rrh

// rrh_di.hip -- ReSTIR DI (restir_di.slang: reservoirs :57-72, p_hat :152-178,
// cs_initial :186-262, cs_temporal :266-335, cs_spatial_shade :339-472; host order
// ReSTIRDIPass::execute, src/passes/restir_di/ReSTIRDIPass.cpp:306-447).
//
// One thread per pixel in 8x8 tiles (one wave64 per tile).  Reservoirs are the
// reference's RGBA32F (light, w_sum, W, M) plus an R32U plane that carries the
// 16-bit-quantised sample point for triangle lights (zero for point lights).
#include "rrh_internal.h"

using namespace rrh;

namespace {

constexpr uint32_t kEmpty = 0xffffffffu;

__device__ __forceinline__ float src_area(const DevScene& S, uint32_t li) {
    if (li < S.nl) return 1.0f;
    return S.tl[(size_t)(li - S.nl) * kTlF4 + 3].y;
}

// restir_di.slang:152-178 (triangle lights add cos_l / d^2)
__device__ __forceinline__ float p_hat(const DevScene& S, uint32_t li, uint32_t uvq, f3 hp, f3 n) {
    uint32_t N = S.nl + S.ntl;
    if (li >= N) return 0.0f;
    f3 l, irr;
    if (li < S.nl) {
        const rrh_light& L = S.lights[li];
        if (L.type == 0) {
            f3 to = sub(ld3(L.position), hp);
            float d2 = fmaxf(dot(to, to), 1e-6f);
            l = normalize(to);
            irr = divs(scale(ld3(L.emission), L.intensity), d2);
        } else {
            l = normalize(neg(ld3(L.direction)));
            irr = scale(ld3(L.emission), L.intensity);
        }
    } else {
        float u1, u2;
        dequant_uv(uvq, u1, u2);
        TriSample ts = tri_light_point(S, li - S.nl, u1, u2);
        f3 to = sub(ts.p, hp);
        float d2 = fmaxf(dot(to, to), 1e-6f);
        l = normalize(to);
        float cosl = fabsf(dot(ts.nl, l));
        irr = scale(ts.Le, cosl / d2);
    }
    float ndl = fmaxf(dot(n, l), 0.0f);
    return ndl * lum709(irr);
}

__device__ __forceinline__ float4 pack(uint32_t sel, float wsum, float W, uint32_t M) {
    return make_float4(__uint_as_float(sel), wsum, W, __uint_as_float(M));
}

template <int STK>
__global__ void __launch_bounds__(64) k_di_initial(DevScene S, uint32_t frame_seed, uint32_t M, float4* __restrict__ gpos,
                                                   float4* __restrict__ gnrm, float4* __restrict__ res,
                                                   uint32_t* __restrict__ res_uv, bool remap) {
    __shared__ int s_stack[StackSel<STK>::kLds * 64];
    const rrh_camera& cam = *S.cam;
    uint32_t W = cam.screen_width, H = cam.screen_height;
    uint32_t tile = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y, remap);
    uint32_t px = (tile % gridDim.x) * 8 + (threadIdx.x & 7u), py = (tile / gridDim.x) * 8 + (threadIdx.x >> 3);
    if (px >= W || py >= H) return;
    size_t i = (size_t)py * W + px;
    uint32_t seed = make_seed(px, py, frame_seed);
    f3 d = primary_dir(cam, px, py, seed);
    f3 o = mk(cam.position[0], cam.position[1], cam.position[2]);
    auto st = make_stack<STK>(s_stack);
    Hit h = trace_closest_s<false>(S, o, d, 0.001f, 1e9f, st);
    res_uv[i] = 0u;
    if (h.prim == ~0u) {
        gpos[i] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
        gnrm[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        res[i] = pack(kEmpty, 0.0f, 0.0f, 0u);
        return;
    }
    HitInfo hi = fetch_hit(S, h.prim, h.b1, h.b2);
    f3 n = hi.normal;
    if (dot(n, neg(d)) < 0.0f) n = neg(n);
    gpos[i] = make_float4(hi.pos.x, hi.pos.y, hi.pos.z, h.t);
    gnrm[i] = make_float4(n.x, n.y, n.z, __uint_as_float(hi.mat));
    uint32_t N = S.nl + S.ntl;
    if (N == 0) { res[i] = pack(kEmpty, 0.0f, 0.0f, 0u); return; }
    uint32_t sel = kEmpty, sel_uv = 0, Mc = 0;
    float wsum = 0.0f, sel_ph = 0.0f;  // sel_ph: p_hat of the kept candidate (the same value a re-evaluation gives)
    for (uint32_t k = 0; k < M; ++k) {
        uint32_t cand = pick_index(rnd(seed), N);
        uint32_t q = 0;
        if (cand >= S.nl) {
            float a = rnd(seed);
            float b = rnd(seed);
            q = quant_uv(a, b);
        }
        float ph = p_hat(S, cand, q, hi.pos, n);
        float wi = ph * (float)N * src_area(S, cand);
        wsum += wi;
        Mc++;
        if (wsum > 0.0f && rnd(seed) < wi / wsum) { sel = cand; sel_uv = q; sel_ph = ph; }
    }
    float Wf = 0.0f;
    if (sel != kEmpty) Wf = (sel_ph > 1e-10f) ? (wsum / ((float)Mc * sel_ph)) : 0.0f;
    res[i] = pack(sel, wsum, Wf, Mc);
    res_uv[i] = sel_uv;
}

__global__ void __launch_bounds__(64) k_di_temporal(DevScene S, uint32_t frame_seed, uint32_t M, const float4* __restrict__ gpos,
                                                    const float4* __restrict__ gnrm, float4* __restrict__ rc,
                                                    uint32_t* __restrict__ rc_uv, const float4* __restrict__ rp,
                                                    const uint32_t* __restrict__ rp_uv, bool remap) {
    const rrh_camera& cam = *S.cam;
    uint32_t W = cam.screen_width, H = cam.screen_height;
    uint32_t tile = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y, remap);
    uint32_t px = (tile % gridDim.x) * 8 + (threadIdx.x & 7u), py = (tile / gridDim.x) * 8 + (threadIdx.x >> 3);
    if (px >= W || py >= H) return;
    size_t i = (size_t)py * W + px;
    float4 gp = gpos[i];
    if (gp.w < 0.0f) return;
    f3 hp = xyz(gp), n = xyz(gnrm[i]);
    float4 c = rc[i];
    uint32_t csel = __float_as_uint(c.x), cM = __float_as_uint(c.w), cuv = rc_uv[i];
    float cw = c.y;
    float4 p = rp[i];
    uint32_t psel = __float_as_uint(p.x), pM = __float_as_uint(p.w), puv = rp_uv[i];
    float pW = p.z;
    if (psel == kEmpty || pM == 0u) return;
    uint32_t cap = M * 20u;
    cap = cap < 1u ? 1u : (cap > 8192u ? 8192u : cap);
    pM = pM < cap ? pM : cap;
    float php = p_hat(S, psel, puv, hp, n);
    float contrib = php * pW * (float)pM;
    cw += contrib;
    uint32_t nM = cM + pM;
    if (nM > cap) { float sc = (float)cap / (float)nM; cw *= sc; cM = cap; }
    else cM = nM;
    uint32_t seed = make_seed(px, py, frame_seed + 1u);
    if (contrib > 0.0f && rnd(seed) < contrib / cw) { csel = psel; cuv = puv; }
    float Wn = 0.0f;
    if (csel != kEmpty) {
        float ph = p_hat(S, csel, cuv, hp, n);
        Wn = (ph > 1e-10f) ? (cw / ((float)cM * ph)) : 0.0f;
        Wn = fminf(Wn, 1e4f);
    }
    rc[i] = pack(csel, cw, Wn, cM);
    rc_uv[i] = cuv;
}

// What a spatial neighbour reads of a pixel, in one 16-byte record: selected light (kEmpty for
// a miss, an empty reservoir or M = 0), W, M and the sample point; one cache line per
// neighbour instead of one in each of the position, reservoir and sample planes
__global__ void __launch_bounds__(256) k_di_pack(const float4* __restrict__ gpos, const float4* __restrict__ res,
                                                 const uint32_t* __restrict__ res_uv, uint4* __restrict__ pk, uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float4 r = res[i];
    uint32_t sel = __float_as_uint(r.x);
    const uint32_t M = __float_as_uint(r.w);
    if (gpos[i].w < 0.0f || M == 0u) sel = kEmpty;
    pk[i] = make_uint4(sel, __float_as_uint(r.z), M, res_uv[i]);
}

template <int STK>
__global__ void __launch_bounds__(64) k_di_spatial_shade(DevScene S, rrh_di_params P, const float4* __restrict__ gpos,
                                                         const float4* __restrict__ gnrm, const float4* __restrict__ res,
                                                         const uint32_t* __restrict__ res_uv, const uint4* __restrict__ pk,
                                                         float4* __restrict__ out, uint32_t* __restrict__ shadow_count,
                                                         bool remap) {
    __shared__ int s_stack[StackSel<STK>::kLds * 64];
    const rrh_camera& cam = *S.cam;
    uint32_t W = cam.screen_width, H = cam.screen_height;
    uint32_t tile = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y, remap);
    uint32_t px = (tile % gridDim.x) * 8 + (threadIdx.x & 7u), py = (tile / gridDim.x) * 8 + (threadIdx.x >> 3);
    bool traced = false;
    if (px < W && py < H) {
        size_t i = (size_t)py * W + px;
        uint32_t N = S.nl + S.ntl;
        float4 gp = gpos[i];
        if (gp.w < 0.0f) {
            f3 sk = sky(center_dir(cam, px, py));
            out[i] = make_float4(sk.x, sk.y, sk.z, 1.0f);
        } else {
            f3 hp = xyz(gp);
            float4 gn = gnrm[i];
            f3 n = xyz(gn);
            Mat m = load_mat(S, __float_as_uint(gn.w));
            if (is_emissive(m)) {
                out[i] = make_float4(m.emissive.x, m.emissive.y, m.emissive.z, 1.0f);
            } else {
                float4 r = res[i];
                uint32_t rsel = __float_as_uint(r.x), rM = __float_as_uint(r.w), ruv = res_uv[i];
                float rw = r.y, rW = r.z;
                if (P.spatial_enable && P.spatial_num_neighbors > 0 && N > 0) {
                    uint32_t seed = make_seed(px, py, P.frame_seed + 2u);
                    for (uint32_t k = 0; k < P.spatial_num_neighbors; ++k) {
                        float angle = rnd(seed) * 6.28318530717958f;
                        float radius = rnd(seed) * (float)P.spatial_radius;
                        float sa, ca;
                        sincos_rr(angle, &sa, &ca);
                        int nx = (int)px + (int)(radius * ca), ny = (int)py + (int)(radius * sa);
                        nx = nx < 0 ? 0 : (nx > (int)W - 1 ? (int)W - 1 : nx);
                        ny = ny < 0 ? 0 : (ny > (int)H - 1 ? (int)H - 1 : ny);
                        size_t j = (size_t)ny * W + (size_t)nx;
                        const uint4 nr = pk[j];
                        const uint32_t nsel = nr.x, nM = nr.z, nuv = nr.w;
                        if (nsel == kEmpty) continue;  // a miss, an empty reservoir or M = 0
                        float ph = p_hat(S, nsel, nuv, hp, n);
                        float contrib = ph * __uint_as_float(nr.y) * (float)nM;
                        rw += contrib;
                        rM += nM;
                        if (contrib > 0.0f && rnd(seed) < contrib / rw) { rsel = nsel; ruv = nuv; }
                    }
                    if (rsel != kEmpty) {
                        float ph = p_hat(S, rsel, ruv, hp, n);
                        rW = (ph > 1e-10f) ? (rw / ((float)rM * ph)) : 0.0f;
                    }
                }
                f3 oc = mk(0, 0, 0);
                if (rsel != kEmpty && rW > 0.0f && N > 0) {
                    f3 l, irr;
                    float dist;
                    if (rsel < S.nl) {
                        const rrh_light& L = S.lights[rsel];
                        if (L.type == 0) {
                            f3 to = sub(ld3(L.position), hp);
                            dist = length(to);
                            l = divs(to, fmaxf(dist, 1e-6f));
                            irr = divs(scale(ld3(L.emission), L.intensity), fmaxf(dist * dist, 1e-6f));
                        } else {
                            l = normalize(neg(ld3(L.direction)));
                            dist = 1e9f;
                            irr = scale(ld3(L.emission), L.intensity);
                        }
                    } else {
                        float u1, u2;
                        dequant_uv(ruv, u1, u2);
                        TriSample ts = tri_light_point(S, rsel - S.nl, u1, u2);
                        f3 to = sub(ts.p, hp);
                        dist = length(to);
                        l = divs(to, fmaxf(dist, 1e-6f));
                        float cosl = fabsf(dot(ts.nl, l));
                        irr = scale(ts.Le, cosl / fmaxf(dist * dist, 1e-6f));
                    }
                    float ndl = fmaxf(dot(n, l), 0.0f);
                    if (ndl > 0.0f) {
                        traced = true;
                        auto st = make_stack<STK>(s_stack);
                        if (!trace_occluded_s(S, add(hp, scale(n, 1e-4f)), l, 0.001f, dist - 0.01f, st)) {
                            f3 wo = normalize(sub(ld3(cam.position), hp));
                            f3 f = mat_eval(m, n, wo, l);
                            oc = scale(scale(mul(f, irr), ndl), rW);
                        }
                    }
                }
                out[i] = make_float4(oc.x, oc.y, oc.z, 1.0f);
            }
        }
    }
    wave_push(shadow_count, traced);
}

__global__ void k_di_totals(const uint32_t* __restrict__ shadow_count, uint32_t npix, unsigned long long* __restrict__ tot) {
    if (threadIdx.x != 0) return;
    tot[0] += npix;
    tot[2] += *shadow_count;
}

bool same_extent(const rrh_image* a, const rrh_image* b) { return a && b && a->w == b->w && a->h == b->h; }

}  // namespace

extern "C" {

int rrh_di_initial(rrh_ctx* c, rrh_scene* s, const rrh_di_params* p, rrh_image* gpos, rrh_image* gnrm, rrh_image* res,
                   rrh_image* res_uv) {
    RRH_ARG(c && s && p && gpos && gnrm && res && res_uv, "rrh_di_initial: null argument");
    RRH_ARG(s->cam_set, "rrh_di_initial: camera not set");
    uint32_t W = s->cam_host.screen_width, H = s->cam_host.screen_height;
    RRH_ARG(gpos->w == W && gpos->h == H, "rrh_di_initial: image extent differs from the camera");
    RRH_ARG(same_extent(gpos, gnrm) && same_extent(gpos, res) && same_extent(gpos, res_uv), "rrh_di_initial: extent mismatch");
    RRH_ARG(gpos->fmt == RRH_FMT_RGBA32F && gnrm->fmt == RRH_FMT_RGBA32F && res->fmt == RRH_FMT_RGBA32F &&
                res_uv->fmt == RRH_FMT_R32U, "rrh_di_initial: formats RGBA32F x3 + R32U");
    RRH_ARG(p->num_initial_candidates <= 4096, "rrh_di_initial: num_initial_candidates above 4096");
    const int sk = stack_for(c, s->stack_need());
    dim3 grid((W + 7) / 8, (H + 7) / 8);
    if (sk == RRH_STACK_LDS32)
        hipLaunchKernelGGL(k_di_initial<RRH_STACK_LDS32>, grid, dim3(64), 0, c->stream, s->dev(), p->frame_seed,
                           p->num_initial_candidates, (float4*)gpos->p, (float4*)gnrm->p, (float4*)res->p,
                           (uint32_t*)res_uv->p, c->xcd_remap);
    else if (sk == RRH_STACK_HYBRID16)
        hipLaunchKernelGGL(k_di_initial<RRH_STACK_HYBRID16>, grid, dim3(64), 0, c->stream, s->dev(), p->frame_seed,
                           p->num_initial_candidates, (float4*)gpos->p, (float4*)gnrm->p, (float4*)res->p,
                           (uint32_t*)res_uv->p, c->xcd_remap);
    else if (sk == RRH_STACK_HYBRID32)
        hipLaunchKernelGGL(k_di_initial<RRH_STACK_HYBRID32>, grid, dim3(64), 0, c->stream, s->dev(), p->frame_seed,
                           p->num_initial_candidates, (float4*)gpos->p, (float4*)gnrm->p, (float4*)res->p,
                           (uint32_t*)res_uv->p, c->xcd_remap);
    else
        hipLaunchKernelGGL(k_di_initial<RRH_STACK_PRIVATE>, grid, dim3(64), 0, c->stream, s->dev(), p->frame_seed,
                           p->num_initial_candidates, (float4*)gpos->p, (float4*)gnrm->p, (float4*)res->p,
                           (uint32_t*)res_uv->p, c->xcd_remap);
    RRH_LAUNCH_CHECK();
    return RRH_OK;
}

int rrh_di_temporal(rrh_ctx* c, rrh_scene* s, const rrh_di_params* p, const rrh_image* gpos, const rrh_image* gnrm,
                    rrh_image* rc, rrh_image* rc_uv, const rrh_image* rp, const rrh_image* rp_uv) {
    RRH_ARG(c && s && p && gpos && gnrm && rc && rc_uv && rp && rp_uv, "rrh_di_temporal: null argument");
    RRH_ARG(s->cam_set, "rrh_di_temporal: camera not set");
    if (!p->temporal_enable) return RRH_OK;
    uint32_t W = s->cam_host.screen_width, H = s->cam_host.screen_height;
    RRH_ARG(gpos->w == W && gpos->h == H, "rrh_di_temporal: image extent differs from the camera");
    RRH_ARG(same_extent(gpos, gnrm) && same_extent(gpos, rc) && same_extent(gpos, rc_uv) && same_extent(gpos, rp) &&
                same_extent(gpos, rp_uv), "rrh_di_temporal: extent mismatch");
    hipLaunchKernelGGL(k_di_temporal, dim3((W + 7) / 8, (H + 7) / 8), dim3(64), 0, c->stream, s->dev(), p->frame_seed,
                       p->num_initial_candidates, (const float4*)gpos->p, (const float4*)gnrm->p, (float4*)rc->p,
                       (uint32_t*)rc_uv->p, (const float4*)rp->p, (const uint32_t*)rp_uv->p, c->xcd_remap);
    RRH_LAUNCH_CHECK();
    return RRH_OK;
}

int rrh_di_spatial_shade(rrh_ctx* c, rrh_scene* s, const rrh_di_params* p, const rrh_image* gpos,
                         const rrh_image* gnrm, const rrh_image* res, const rrh_image* res_uv, rrh_image* out) {
    RRH_ARG(c && s && p && gpos && gnrm && res && res_uv && out, "rrh_di_spatial_shade: null argument");
    RRH_ARG(s->cam_set, "rrh_di_spatial_shade: camera not set");
    uint32_t W = s->cam_host.screen_width, H = s->cam_host.screen_height;
    RRH_ARG(gpos->w == W && gpos->h == H, "rrh_di_spatial_shade: image extent differs from the camera");
    RRH_ARG(same_extent(gpos, gnrm) && same_extent(gpos, res) && same_extent(gpos, res_uv) && same_extent(gpos, out),
            "rrh_di_spatial_shade: extent mismatch");
    RRH_ARG(out->fmt == RRH_FMT_RGBA32F, "rrh_di_spatial_shade: output must be RGBA32F");
    int rc = c->di_counter.ensure(16);
    if (rc) return rc;
    RRH_HIP(hipMemsetAsync(c->di_counter.p, 0, 4, c->stream));
    const uint32_t npix = W * H;
    if ((rc = c->di_pack.ensure((size_t)npix * sizeof(uint4)))) return rc;
    const uint4* pk = (const uint4*)c->di_pack.p;
    hipLaunchKernelGGL(k_di_pack, dim3((npix + 255) / 256), dim3(256), 0, c->stream, (const float4*)gpos->p,
                       (const float4*)res->p, (const uint32_t*)res_uv->p, (uint4*)c->di_pack.p, npix);
    RRH_LAUNCH_CHECK();
    const int sk = stack_for(c, s->stack_need());
    dim3 grid((W + 7) / 8, (H + 7) / 8);
    auto go = [&](auto k) {
        hipLaunchKernelGGL(k, grid, dim3(64), 0, c->stream, s->dev(), *p, (const float4*)gpos->p, (const float4*)gnrm->p,
                           (const float4*)res->p, (const uint32_t*)res_uv->p, pk, (float4*)out->p,
                           (uint32_t*)c->di_counter.p, c->xcd_remap);
    };
    if (sk == RRH_STACK_LDS32) go(k_di_spatial_shade<RRH_STACK_LDS32>);
    else if (sk == RRH_STACK_HYBRID16) go(k_di_spatial_shade<RRH_STACK_HYBRID16>);
    else if (sk == RRH_STACK_HYBRID32) go(k_di_spatial_shade<RRH_STACK_HYBRID32>);
    else go(k_di_spatial_shade<RRH_STACK_PRIVATE>);
    RRH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_di_totals, dim3(1), dim3(64), 0, c->stream, (const uint32_t*)c->di_counter.p, W * H,
                       (unsigned long long*)c->totals.p);
    RRH_LAUNCH_CHECK();
    c->last_npix = W * H;
    return RRH_OK;
}

int rrh_di_ray_counts(rrh_ctx* c, rrh_ray_counts* out) {
    RRH_ARG(c && out, "rrh_di_ray_counts: null argument");
    out->primary = out->extension = out->shadow = 0;
    if (!c->di_counter.p) return RRH_OK;
    uint32_t h = 0;
    RRH_HIP(hipMemcpyAsync(&h, c->di_counter.p, 4, hipMemcpyDeviceToHost, c->stream));
    RRH_HIP(hipStreamSynchronize(c->stream));
    out->shadow = h;
    out->primary = c->last_npix;
    return RRH_OK;
}

}  // extern "C"

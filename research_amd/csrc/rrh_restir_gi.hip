// rrh_restir_gi.hip -- ReSTIR GI: one-bounce indirect reuse with the reconnection shift
// (restir_gi.slang: reservoir planes :204-242, evaluate_indirect_radiance :487-508,
// sample_primary_pdf :510-522, target :524-537, evaluate_sample_radiance :539-551, cs_initial
// :566-684, cs_temporal :686-847, cs_spatial_shade :849-1013; host order ReSTIRGIPass::execute,
// src/passes/restir_gi/ReSTIRGIPass.cpp:428-651).
//
// The single-bounce sibling of ReSTIR PT (rrh_restir_pt.hip) and built the same way: one thread
// per pixel, the initial stage a per-lane state machine around one traversal call in persistent
// waves (candidate segment, then its light sample's shadow ray; rays two-sided unless
// RRH_OPT_RESTIR_CULL selects the culling kernels, as in rrh_restir_pt.hip), the reuse stages
// one pass each with one visibility ray.  The shared pieces (light sampling, Jacobian,
// visibility, history search, neighbour spiral) are in rrh_reuse.h.  Differences from ReSTIR
// PT that the reference has and that are kept: cosine candidates (no guiding, no selection
// draw), the source pdf taken from the hit position (cos / pi toward it), the kept sample's
// normal stored as interpolated (not flipped), three reservoir planes with M = 0 as the empty
// mark, and brdf * radiance * cos as the order of the target's product.
#include "rrh_internal.h"
#include "rrh_reuse.h"

using namespace rrh;
using namespace rrh::reuse;

namespace {

// three RGBA32F planes (restir_gi.slang:204-242)
struct Res3 {
    float4* pos;   // sample position, w_sum
    float4* rad;   // sample radiance, W
    float4* meta;  // sample normal, asfloat(M)
};
struct GiRes {
    f3 pos, rad, n;
    float w_sum, W;
    uint32_t M;
};
RRH_D GiRes load_gi(const Res3& R, size_t i) {
    const float4 p = R.pos[i], r = R.rad[i], m = R.meta[i];
    GiRes s;
    s.pos = xyz(p); s.w_sum = p.w;
    s.rad = xyz(r); s.W = r.w;
    s.n = xyz(m); s.M = __float_as_uint(m.w);
    return s;
}
RRH_D void store_gi(const Res3& R, size_t i, const GiRes& s) {
    R.pos[i] = make_float4(s.pos.x, s.pos.y, s.pos.z, s.w_sum);
    R.rad[i] = make_float4(s.rad.x, s.rad.y, s.rad.z, s.W);
    R.meta[i] = make_float4(s.n.x, s.n.y, s.n.z, __uint_as_float(s.M));
}
RRH_D void store_gi_empty(const Res3& R, size_t i) {
    const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    R.pos[i] = z; R.rad[i] = z; R.meta[i] = z;
}

// evaluate_indirect_radiance / compute_target_function, restir_gi.slang:487-537
RRH_D f3 indirect(f3 pp, f3 pn, const Mat& pm, f3 sp, f3 srad, f3 cam) {
    const f3 to = sub(sp, pp);
    const float d2 = dot(to, to);
    if (d2 <= 1e-8f) return mk(0.0f, 0.0f, 0.0f);
    const f3 wi = scale(to, 1.0f / sqrtf(d2));
    const float cx = fmaxf(dot(pn, wi), 0.0f);
    if (cx <= 0.0f) return mk(0.0f, 0.0f, 0.0f);
    const f3 wo = normalize(sub(cam, pp));
    return scale(mul(mat_eval(pm, pn, wo, wi), srad), cx);
}
RRH_D float gi_target(f3 pp, f3 pn, const Mat& pm, f3 sp, f3 srad, f3 cam) {
    return lum709(indirect(pp, pn, pm, sp, srad, cam));
}
// sample_primary_pdf, restir_gi.slang:510-522: the cosine pdf toward the hit
RRH_D float primary_pdf(f3 pp, f3 pn, f3 sp) {
    const f3 d = sub(sp, pp);
    const float d2 = dot(d, d);
    if (d2 <= 1e-8f) return 0.0f;
    const f3 wi = scale(d, 1.0f / sqrtf(d2));
    const float cx = fmaxf(dot(pn, wi), 0.0f);
    return fmaxf(cx * K_INV_PI, 0.0f);
}

// ---------------------------------------------------------------------------
// cs_initial, restir_gi.slang:566-684 (state machine and persistent waves as in ReSTIR PT)
// ---------------------------------------------------------------------------
enum : uint32_t { gNext, gToHit, gShadow, gFinish, gDone };

template <int KS>
__global__ void __launch_bounds__(64, 4) k_rgi_initial(DevScene S, rrh_rgi_params P, Frame F, Res3 R,
                                                       float4* __restrict__ hpos, float4* __restrict__ hnrm,
                                                       uint32_t* __restrict__ counter, uint32_t tiles_x,
                                                       uint32_t npad) {
    constexpr int STK = KS & ~kCull;
    constexpr bool CULL = (KS & kCull) != 0;
    __shared__ int s_stack[StackSel<STK>::kLds * 64];
    uint32_t rays = 0, shadow_rays = 0;
    uint32_t* next_pixel = counter + 3;
    const f3 cp = ld3(S.cam->position);
    const float bias = P.ray_bias;
    const bool hit_emission = S.ntl == 0;
    size_t i = 0;
    f3 pp = mk(0.0f, 0.0f, 0.0f), pn = pp;
    uint32_t pmi = 0, seed = 0;
    uint32_t state = gDone;
    bool exhausted = false;
    float w_sum = 0.0f;
    uint32_t Mc = 0, c = 0;
    f3 dir = pp, sp = pp, sn = pp, srad = pp, nee = pp;
    f3 ro = pp, rd = pp;
    float rtmax = 0.0f;
    bool first_hit = false;
    Hit h;
    h.t = 0.0f; h.prim = ~0u; h.b1 = h.b2 = 0.0f;
    auto st = make_stack<STK>(s_stack);
    while (true) {
        bool want_ray = false;
        while (true) {
            while (state != gDone && !want_ray) {
                switch (state) {
                case gNext: {
                    if (c >= P.num_initial_candidates) {
                        bool stored = false;
                        if (Mc > 0) {
                            GiRes r;
                            r.pos = xyz(R.pos[i]); r.rad = xyz(R.rad[i]); r.n = xyz(R.meta[i]);
                            const float ph = gi_target(pp, pn, load_mat(S, pmi), r.pos, r.rad, cp);
                            if (ph > 1e-10f) {
                                r.w_sum = w_sum;
                                r.W = w_sum / ((float)Mc * ph);
                                r.M = Mc;
                                store_gi(R, i, r);
                                stored = true;
                            }
                        }
                        if (!stored) store_gi_empty(R, i);
                        state = gDone;
                        break;
                    }
                    ++c;
                    const float u0 = rnd(seed), u1 = rnd(seed);
                    float pdf_dir;
                    const f3 local = cosine_hemisphere(u0, u1, pdf_dir);
                    f3 T, B;
                    build_onb(pn, T, B);
                    dir = to_world(local, T, B, pn);
                    if (dot(pn, dir) <= 0.0f) break;  // dropped, not counted
                    ro = add(pp, scale(pn, bias));
                    rd = dir;
                    rtmax = 1e9f;
                    first_hit = false;
                    rays++;
                    state = gToHit;
                    want_ray = true;
                    break;
                }
                case gToHit: {
                    if (h.prim == ~0u) { state = gNext; break; }  // escaped: dropped, not counted
                    const HitInfo hi = fetch_hit(S, h.prim, h.b1, h.b2);
                    const f3 wo = neg(dir);
                    f3 n = hi.normal;
                    if (dot(n, wo) < 0.0f) n = neg(n);
                    const Mat m = load_mat(S, hi.mat);
                    sp = hi.pos;
                    sn = hi.normal;
                    srad = hit_emission ? m.emissive : mk(0.0f, 0.0f, 0.0f);
                    if (light_sample(S, sp, n, wo, m, bias, seed, nee, ro, rd, rtmax)) {
                        first_hit = true;
                        shadow_rays++;
                        state = gShadow;
                        want_ray = true;
                    } else {
                        srad = add(srad, nee);
                        state = gFinish;
                    }
                    break;
                }
                case gShadow:
                    srad = add(srad, h.prim == ~0u ? nee : mk(0.0f, 0.0f, 0.0f));
                    state = gFinish;
                    break;
                case gFinish: {
                    state = gNext;
                    const float tgt = gi_target(pp, pn, load_mat(S, pmi), sp, srad, cp);
                    if (tgt <= 0.0f) break;
                    const float pdf = primary_pdf(pp, pn, sp);
                    if (pdf <= 1e-6f) break;
                    const float w = tgt / pdf;
                    w_sum += w;
                    Mc++;
                    if (w_sum > 0.0f && rnd(seed) < w / w_sum) {
                        R.pos[i] = make_float4(sp.x, sp.y, sp.z, 0.0f);
                        R.rad[i] = make_float4(srad.x, srad.y, srad.z, 0.0f);
                        R.meta[i] = make_float4(sn.x, sn.y, sn.z, 0.0f);
                    }
                    break;
                }
                default:
                    state = gDone;
                    break;
                }
            }
            const bool need = state == gDone && !exhausted;
            if (!__any(need)) break;
            const uint32_t slot = wave_push(next_pixel, need);
            if (need) {
                if (slot >= npad) {
                    exhausted = true;
                } else {
                    const uint32_t tile = slot >> 6, l = slot & 63u;
                    const uint32_t px = (tile % tiles_x) * 8 + (l & 7u), py = (tile / tiles_x) * 8 + (l >> 3);
                    if (px < F.W && py < F.H) {
                        i = (size_t)py * F.W + px;
                        if (load_surface(F.gpos, F.gnrm, i, pp, pn, pmi)) {
                            hpos[i] = F.gpos[i];
                            hnrm[i] = F.gnrm[i];
                            seed = make_seed(px, py, P.frame_seed);
                            w_sum = 0.0f;
                            Mc = 0;
                            c = 0;
                            state = gNext;
                        } else {
                            hpos[i] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
                            hnrm[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                            store_gi_empty(R, i);
                        }
                    }
                }
            }
        }
        if (!__any(want_ray)) break;
        if (want_ray) {
            if (S.speculative)
                h = trace_closest4spec_s<false, true, CULL>(S, ro, rd, bias, rtmax, st, nullptr, nullptr, first_hit);
            else
                h = trace_closest4_s<false, true, CULL>(S, ro, rd, bias, rtmax, st, nullptr, nullptr, first_hit);
        }
    }
    count_rays(counter, rays);
    count_rays(counter + 1, shadow_rays);
}

// ---------------------------------------------------------------------------
// cs_temporal, restir_gi.slang:686-847
// ---------------------------------------------------------------------------
template <int KS>
__global__ void __launch_bounds__(64) k_rgi_temporal(DevScene S, rrh_rgi_params P, Frame F, Res3 R, Res3 Rp,
                                                     History Hp, uint32_t* __restrict__ counter, bool remap) {
    constexpr int STK = KS & ~kCull;
    constexpr bool CULL = (KS & kCull) != 0;
    __shared__ int s_stack[StackSel<STK>::kLds * 64];
    uint32_t px, py, rays = 0;
    if (tile_pixel(px, py, F.W, F.H, remap)) {
        const size_t i = (size_t)py * F.W + px;
        f3 pp, pn;
        uint32_t pmi;
        const rrh_camera& cam = *S.cam;
        const f3 cp = ld3(cam.position);
        size_t hj;
        f3 hp;
        if (load_surface(F.gpos, F.gnrm, i, pp, pn, pmi) && find_history(cam, Hp, F.W, F.H, pp, pn, pmi, cp, hj, hp)) {
            GiRes cur = load_gi(R, i);
            GiRes prv = load_gi(Rp, hj);
            if (prv.M != 0) {
                uint32_t cap = 20u * P.num_initial_candidates;
                cap = cap < 1u ? 1u : (cap > 8192u ? 8192u : cap);
                prv.M = prv.M < cap ? prv.M : cap;
                const float jac = jacobian(hp, pp, prv.pos, prv.n);
                bool go = jac > 0.0f;
                if (go) {
                    auto st = make_stack<STK>(s_stack);
                    rays++;
                    go = visible<CULL>(S, pp, pn, prv.pos, P.ray_bias, st);
                }
                if (go) {
                    const Mat pm = load_mat(S, pmi);
                    const float ph = gi_target(pp, pn, pm, prv.pos, prv.rad, cp);
                    const float contrib = ph * fminf(prv.W, 50.0f) * (float)prv.M * jac;
                    if (contrib > 0.0f) {
                        cur.w_sum += contrib;
                        uint32_t seed = make_seed(px, py, P.frame_seed + 1u);
                        if (cur.w_sum > 0.0f && rnd(seed) < contrib / cur.w_sum) {
                            cur.pos = prv.pos; cur.rad = prv.rad; cur.n = prv.n;
                        }
                        const uint32_t nM = cur.M + prv.M;
                        if (nM > cap) {
                            cur.w_sum *= (float)cap / (float)nM;
                            cur.M = cap;
                        } else {
                            cur.M = nM;
                        }
                        float Wn = 0.0f;
                        if (cur.M > 0) {
                            const float py_ = gi_target(pp, pn, pm, cur.pos, cur.rad, cp);
                            Wn = (py_ > 1e-10f) ? (cur.w_sum / ((float)cur.M * py_)) : 0.0f;
                            Wn = fminf(Wn, 50.0f);
                        }
                        cur.W = Wn;
                        store_gi(R, i, cur);
                    }
                }
            }
        }
    }
    count_rays(counter, rays);
}

// ---------------------------------------------------------------------------
// cs_spatial_shade, restir_gi.slang:849-1013
// ---------------------------------------------------------------------------
template <int KS>
__global__ void __launch_bounds__(64) k_rgi_spatial_shade(DevScene S, rrh_rgi_params P, Frame F, Res3 R,
                                                          const float4* __restrict__ direct, float4* __restrict__ out,
                                                          uint32_t* __restrict__ counter, bool remap) {
    constexpr int STK = KS & ~kCull;
    constexpr bool CULL = (KS & kCull) != 0;
    __shared__ int s_stack[StackSel<STK>::kLds * 64];
    uint32_t px, py, rays = 0;
    if (tile_pixel(px, py, F.W, F.H, remap)) {
        const size_t i = (size_t)py * F.W + px;
        const rrh_camera& cam = *S.cam;
        const f3 cp = ld3(cam.position);
        f3 L = mk(0.0f, 0.0f, 0.0f);
        if (P.include_direct_lighting != 0 && direct != nullptr) L = xyz(direct[i]);
        f3 pp, pn;
        uint32_t pmi;
        if (!load_surface(F.gpos, F.gnrm, i, pp, pn, pmi)) {
            if (P.include_direct_lighting == 0) L = add(L, sky(center_dir(cam, px, py)));
        } else {
            const Mat pm = load_mat(S, pmi);
            if (is_emissive(pm)) {
                L = add(L, pm.emissive);
            } else {
                GiRes res = load_gi(R, i);
                if (P.spatial_enable && P.spatial_num_neighbors > 0) {
                    const float phase = spiral_phase(px, py);
                    const float dist_scale = fmaxf(length(sub(pp, cp)), 1.0f);
                    for (uint32_t k = 0; k < P.spatial_num_neighbors; ++k) {
                        size_t j;
                        f3 q;
                        if (!neighbour(k, P.spatial_num_neighbors, P.spatial_radius, phase, px, py, F.W, F.H, j)) continue;
                        if (!neighbour_surface(F, j, pp, pn, dist_scale, q)) continue;
                        const GiRes nb = load_gi(R, j);
                        if (nb.M == 0) continue;
                        const float jac = jacobian(q, pp, nb.pos, nb.n);
                        if (jac <= 0.0f) continue;
                        const float contrib = gi_target(pp, pn, pm, nb.pos, nb.rad, cp) * nb.W * (float)nb.M * jac;
                        if (contrib <= 0.0f) continue;
                        res.w_sum += contrib;
                        res.M += nb.M;
                        const float accept = fracf(phase + 0.754877666f * (float)(k + 1u));
                        if (res.w_sum > 0.0f && accept < contrib / res.w_sum) {
                            res.pos = nb.pos; res.rad = nb.rad; res.n = nb.n;
                        }
                    }
                    if (res.M > 0) {
                        const float ph = gi_target(pp, pn, pm, res.pos, res.rad, cp);
                        res.W = (ph > 1e-10f) ? (res.w_sum / ((float)res.M * ph)) : 0.0f;
                    }
                }
                if (res.M > 0 && res.W > 0.0f) {
                    const f3 ind = scale(indirect(pp, pn, pm, res.pos, res.rad, cp), res.W);
                    if (lum709(ind) > 0.0f) {
                        auto st = make_stack<STK>(s_stack);
                        rays++;
                        if (visible<CULL>(S, pp, pn, res.pos, P.ray_bias, st)) L = add(L, scale(ind, P.indirect_strength));
                    }
                }
            }
        }
        out[i] = make_float4(L.x, L.y, L.z, 1.0f);
    }
    count_rays(counter, rays);
}

__global__ void k_rgi_clear(Res3 R, float4* __restrict__ hpos, float4* __restrict__ hnrm, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    store_gi_empty(R, i);
    hpos[i] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
    hnrm[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// counters: [0] candidate segments, [1] their shadow rays and the temporal visibility rays,
// [2] the spatial stage's visibility rays, [3] the initial stage's next pixel
__global__ void k_rgi_totals(const uint32_t* __restrict__ cnt, unsigned long long* __restrict__ tot) {
    if (threadIdx.x != 0) return;
    tot[1] += cnt[0];
    tot[2] += cnt[1] + cnt[2];
}

// ---- host side -----------------------------------------------------------------------
int res3(rrh_image* const* r, uint32_t W, uint32_t H, Res3& out, const char* who) {
    if (!r) { set_error(std::string(who) + ": null reservoir"); return RRH_E_ARG; }
    for (int k = 0; k < 3; ++k)
        if (!fits(r[k], W, H)) {
            set_error(std::string(who) + ": reservoir planes must be three RGBA32F images of the camera's extent");
            return RRH_E_ARG;
        }
    out = Res3{(float4*)r[0]->p, (float4*)r[1]->p, (float4*)r[2]->p};
    return RRH_OK;
}

int prepare(rrh_ctx* c, rrh_scene* s, const rrh_rgi_params* p, const rrh_image* gpos, const rrh_image* gnrm,
            Frame& F, const char* who) {
    if (!p) { set_error(std::string(who) + ": null argument"); return RRH_E_ARG; }
    int rc = check_frame(c, s, gpos, gnrm, p->num_initial_candidates, p->spatial_num_neighbors, F, who);
    if (rc) return rc;
    return c->rgi_counter.ensure(16);
}

template <int S> struct KInit { static constexpr auto fn = k_rgi_initial<S>; };
template <int S> struct KTemp { static constexpr auto fn = k_rgi_temporal<S>; };
template <int S> struct KSpat { static constexpr auto fn = k_rgi_spatial_shade<S>; };

}  // namespace

extern "C" {

int rrh_rgi_initial(rrh_ctx* c, rrh_scene* s, const rrh_rgi_params* p, const rrh_image* gpos, const rrh_image* gnrm,
                    rrh_image* const res[3], rrh_image* hist_pos, rrh_image* hist_norm) {
    Frame F;
    int rc = prepare(c, s, p, gpos, gnrm, F, "rrh_rgi_initial");
    if (rc) return rc;
    RRH_ARG(fits(hist_pos, F.W, F.H) && fits(hist_norm, F.W, F.H),
            "rrh_rgi_initial: history position / normal must be RGBA32F images of the camera's extent");
    Res3 R;
    if ((rc = res3(res, F.W, F.H, R, "rrh_rgi_initial"))) return rc;
    RRH_HIP(hipMemsetAsync(c->rgi_counter.p, 0, 16, c->stream));
    const int sk = kernel_kind(c, s);
    const uint32_t tiles_x = (F.W + 7) / 8, tiles = tiles_x * ((F.H + 7) / 8);
    uint32_t waves = 0;
    rc = persistent_waves<KInit>(c, sk, tiles, waves);
    if (rc) return rc;
    launch<KInit>(sk, dim3(waves), c->stream, s->dev(), *p, F, R, (float4*)hist_pos->p, (float4*)hist_norm->p,
                  (uint32_t*)c->rgi_counter.p, tiles_x, tiles * 64u);
    RRH_LAUNCH_CHECK();
    return RRH_OK;
}

int rrh_rgi_temporal(rrh_ctx* c, rrh_scene* s, const rrh_rgi_params* p, const rrh_image* gpos, const rrh_image* gnrm,
                     rrh_image* const res[3], rrh_image* const res_prev[3], const rrh_image* hist_prev_pos,
                     const rrh_image* hist_prev_norm) {
    Frame F;
    int rc = prepare(c, s, p, gpos, gnrm, F, "rrh_rgi_temporal");
    if (rc) return rc;
    RRH_ARG(fits(hist_prev_pos, F.W, F.H) && fits(hist_prev_norm, F.W, F.H),
            "rrh_rgi_temporal: history position / normal must be RGBA32F images of the camera's extent");
    Res3 R, Rp;
    if ((rc = res3(res, F.W, F.H, R, "rrh_rgi_temporal"))) return rc;
    if ((rc = res3(res_prev, F.W, F.H, Rp, "rrh_rgi_temporal"))) return rc;
    if (!p->temporal_enable) return RRH_OK;
    launch<KTemp>(kernel_kind(c, s), dim3((F.W + 7) / 8, (F.H + 7) / 8), c->stream, s->dev(), *p, F, R, Rp,
                  History{(const float4*)hist_prev_pos->p, (const float4*)hist_prev_norm->p},
                  (uint32_t*)c->rgi_counter.p + 1, c->xcd_remap);
    RRH_LAUNCH_CHECK();
    return RRH_OK;
}

int rrh_rgi_spatial_shade(rrh_ctx* c, rrh_scene* s, const rrh_rgi_params* p, const rrh_image* gpos,
                          const rrh_image* gnrm, const rrh_image* direct, rrh_image* const res[3], rrh_image* out) {
    Frame F;
    int rc = prepare(c, s, p, gpos, gnrm, F, "rrh_rgi_spatial_shade");
    if (rc) return rc;
    RRH_ARG(fits(out, F.W, F.H), "rrh_rgi_spatial_shade: output must be an RGBA32F image of the camera's extent");
    RRH_ARG(!direct || fits(direct, F.W, F.H),
            "rrh_rgi_spatial_shade: direct lighting must be an RGBA32F image of the camera's extent");
    Res3 R;
    if ((rc = res3(res, F.W, F.H, R, "rrh_rgi_spatial_shade"))) return rc;
    launch<KSpat>(kernel_kind(c, s), dim3((F.W + 7) / 8, (F.H + 7) / 8), c->stream, s->dev(), *p, F, R,
                  direct ? (const float4*)direct->p : (const float4*)nullptr, (float4*)out->p,
                  (uint32_t*)c->rgi_counter.p + 2, c->xcd_remap);
    RRH_LAUNCH_CHECK();
    if (!c->totals.p) {
        if ((rc = c->totals.ensure(32))) return rc;
        RRH_HIP(hipMemsetAsync(c->totals.p, 0, 32, c->stream));
    }
    hipLaunchKernelGGL(k_rgi_totals, dim3(1), dim3(64), 0, c->stream, (const uint32_t*)c->rgi_counter.p,
                       (unsigned long long*)c->totals.p);
    RRH_LAUNCH_CHECK();
    return RRH_OK;
}

int rrh_rgi_clear(rrh_ctx* c, rrh_image* const res[3], rrh_image* hist_pos, rrh_image* hist_norm) {
    RRH_ARG(c && res && hist_pos && hist_norm, "rrh_rgi_clear: null argument");
    const uint32_t W = hist_pos->w, H = hist_pos->h;
    RRH_ARG(fits(hist_pos, W, H) && fits(hist_norm, W, H), "rrh_rgi_clear: history planes must be RGBA32F, same extent");
    Res3 R;
    int rc = res3(res, W, H, R, "rrh_rgi_clear");
    if (rc) return rc;
    const uint32_t n = W * H;
    if (n == 0) return RRH_OK;
    hipLaunchKernelGGL(k_rgi_clear, dim3((n + 255) / 256), dim3(256), 0, c->stream, R, (float4*)hist_pos->p,
                       (float4*)hist_norm->p, n);
    RRH_LAUNCH_CHECK();
    return RRH_OK;
}

int rrh_rgi_ray_counts(rrh_ctx* c, rrh_ray_counts* out) {
    RRH_ARG(c && out, "rrh_rgi_ray_counts: null argument");
    out->primary = out->extension = out->shadow = 0;
    if (!c->rgi_counter.p) return RRH_OK;
    uint32_t h[4] = {0, 0, 0, 0};
    RRH_HIP(hipMemcpyAsync(h, c->rgi_counter.p, 16, hipMemcpyDeviceToHost, c->stream));
    RRH_HIP(hipStreamSynchronize(c->stream));
    out->extension = h[0];
    out->shadow = (uint64_t)h[1] + h[2];
    return RRH_OK;
}

}  // extern "C"

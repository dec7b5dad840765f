// rrh_restir_pt.hip -- ReSTIR PT with the reconnection shift (restir_pt.slang: reservoir
// planes :11-25, trace_ray :98-161, sample_lights :171-222, history and reprojection :273-421,
// reconnection_jacobian :427-447, trace_visibility :449-458, target function :464-489,
// cs_initial :569-771, cs_temporal :775-872, cs_spatial_shade :876-1001; host order
// ReSTIRPTPass::execute, src/passes/restir_pt/ReSTIRPTPass.cpp:529-813).
//
// One thread per pixel in 8x8 tiles (one wave64 per tile), every stage one kernel: the
// initial stage traces num_initial_candidates paths of up to max_bounces segments from the
// primary surface (extension and shadow rays in the same thread, traversal stack in LDS),
// resamples them by their reconnected contribution and keeps one; the temporal stage
// reprojects into the previous frame and reconnects its kept sample; the spatial stage
// reconnects neighbours' samples, shades with one visibility ray and splats the chosen
// direction into the guiding grid.  Rays are two-sided: the shader's RAY_FLAG_CULL_BACK_FACING_
// TRIANGLES is overridden by the reference's CULL_DISABLE instances (AccelStructure.cpp:90);
// RRH_OPT_RESTIR_CULL selects kernels that cull (a template flag, so each kernel keeps one
// traversal).  Rays traverse the 4-wide nodes (quantized; the visibility rays of the reuse
// stages read the fp32 4-wide nodes when the context selects them).
//
// Deviations, each for exactness or for the triangle-light scenes:
//  * 1/sqrtf(x) where the reference writes rsqrt(x), so host and device round alike;
//  * in scenes whose emissive triangles are lights (RRH_SCENE_TRI_LIGHTS), light sampling
//    covers them (area sampling, no MIS: emission is never added at a BSDF hit of the suffix)
//    and the reconnection vertex's own emission is left to the direct lighting pass, which
//    samples the same lights; with point / directional lights only, as in the reference, the
//    reconnection vertex's emission is part of the suffix.
#include "rrh_guide_dev.h"
#include "rrh_internal.h"
#include "rrh_reuse.h"

using namespace rrh;
using namespace rrh::reuse;

namespace {

// four RGBA32F planes (restir_pt.slang:11-15)
struct Res4 {
    float4* pos;  // reconnection position, w_sum
    float4* nrm;  // reconnection normal, asfloat(M)
    float4* rad;  // suffix radiance, W
    float4* pre;  // prefix throughput, valid
};
struct Reservoir {
    f3 rc_pos, rc_n, suffix, prefix;
    float w_sum, W;
    uint32_t M;
    float valid;
};

RRH_D Reservoir load_res(const Res4& R, size_t i) {
    const float4 p = R.pos[i], n = R.nrm[i], r = R.rad[i], e = R.pre[i];
    Reservoir s;
    s.rc_pos = xyz(p); s.w_sum = p.w;
    s.rc_n = xyz(n); s.M = __float_as_uint(n.w);
    s.suffix = xyz(r); s.W = r.w;
    s.prefix = xyz(e); s.valid = e.w;
    return s;
}
RRH_D void store_res(const Res4& R, size_t i, const Reservoir& s) {
    R.pos[i] = make_float4(s.rc_pos.x, s.rc_pos.y, s.rc_pos.z, s.w_sum);
    R.nrm[i] = make_float4(s.rc_n.x, s.rc_n.y, s.rc_n.z, __uint_as_float(s.M));
    R.rad[i] = make_float4(s.suffix.x, s.suffix.y, s.suffix.z, s.W);
    R.pre[i] = make_float4(s.prefix.x, s.prefix.y, s.prefix.z, s.valid);
}
RRH_D void store_empty(const Res4& R, size_t i) {
    const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    R.pos[i] = z; R.nrm[i] = z; R.rad[i] = z; R.pre[i] = z;
}

// evaluate_contribution / target_function, restir_pt.slang:464-489
RRH_D f3 contribution(f3 pp, f3 pn, const Mat& pm, f3 cam, const Reservoir& r) {
    const f3 to = sub(r.rc_pos, pp);
    const float d2 = dot(to, to);
    if (d2 <= 1e-8f) return mk(0.0f, 0.0f, 0.0f);
    const f3 wi = scale(to, 1.0f / sqrtf(d2));
    const float cx = fmaxf(dot(pn, wi), 0.0f);
    if (cx <= 0.0f) return mk(0.0f, 0.0f, 0.0f);
    const f3 wo = normalize(sub(cam, pp));
    const f3 f = mat_eval(pm, pn, wo, wi);
    return mul(mul(scale(f, cx), r.prefix), r.suffix);
}
RRH_D float target(f3 pp, f3 pn, const Mat& pm, f3 cam, const Reservoir& r) {
    return lum709(contribution(pp, pn, pm, cam, r));
}

// ---------------------------------------------------------------------------
// cs_initial, restir_pt.slang:569-771
// ---------------------------------------------------------------------------
// The candidate loop runs as a per-lane state machine around ONE traversal call: each pass of
// the outer loop lets every lane consume the result of its ray in flight and advance to its
// next ray (the candidate's first segment, a light sample's shadow ray, or the next path
// segment), then the whole wave traverses together.  The waves are persistent: a lane that
// finishes its pixel takes the next one from a counter (pixels in 8x8 tile order), so a wave's
// lanes stay busy until the image runs out instead of waiting for the tile's longest pixel.
// A pixel's random numbers and arithmetic follow the reference's nested loops exactly; only
// which lanes traverse together changes.
enum : uint32_t { kNext, kToRc, kShadow, kBounce, kExt, kFinish, kDone };

template <int KS, bool STEPS>
__global__ void __launch_bounds__(64, 4) k_rpt_initial(DevScene S, rrh_rpt_params P, HashGrid G, Frame F, Res4 R,
                                                    float4* __restrict__ hpos, float4* __restrict__ hnrm,
                                                    uint32_t* __restrict__ counter, uint32_t tiles_x,
                                                    uint32_t npad, uint32_t step_visits) {
    constexpr int STK = KS & ~kCull;
    constexpr bool CULL = (KS & kCull) != 0;
    __shared__ int s_stack[StackSel<STK>::kLds * 64];
    uint32_t rays = 0, shadow_rays = 0;
    uint32_t* next_pixel = counter + 3;
    const f3 cp = ld3(S.cam->position);
    const float bias = P.ray_bias;
    const bool rc_emission = S.ntl == 0;
    const float alpha = P.guiding_enable ? P.guiding_alpha : 0.0f;
    const uint32_t max_b = P.max_bounces > 1u ? P.max_bounces : 1u;
    // the pixel (its material is read again where needed: fewer live registers)
    size_t i = 0;
    f3 pp = mk(0.0f, 0.0f, 0.0f), pn = pp;
    uint32_t pmi = 0, seed = 0;
    // the resampling state of the pixel (the kept candidate waits in the reservoir planes) and
    // the candidate in progress (the current vertex's material by index; its outgoing
    // direction is -dir)
    uint32_t state = kDone;
    bool exhausted = false, have_sel = false;
    float w_sum = 0.0f, src_pdf = 0.0f;
    uint32_t Mc = 0, c = 0, b = 0, cur_mi = 0;
    f3 dir = pp;  // the candidate's first direction, then the last segment's
    f3 rc_pos = pp, rc_n = pp, suffix = pp, thp = pp, nee = pp, cur_p = pp, cur_n = pp;
    // the ray in flight
    f3 ro = pp, rd = pp;
    float rtmax = 0.0f;
    bool first_hit = false;
    Hit h;
    h.t = 0.0f; h.prim = ~0u; h.b1 = h.b2 = 0.0f;
    auto st = make_stack<STK>(s_stack);
    // STEPS: a ray in flight is traced in passes of at most step_visits node visits
    // (rrh::trace4q_steps), and a lane whose ray has finished runs its state machine while the
    // others' rays are still in flight: `tracing` marks a ray in flight, (tr, tnode, tsp) and the
    // lane's stack hold where it stands
    constexpr bool steps = STEPS;  // the host picks it with the speculative traversal only
    bool tracing = false;
    int tnode = kTravDone, tsp = 0;
    WRay tr = wray_setup(mk(0.0f, 0.0f, 0.0f), mk(0.0f, 0.0f, 1.0f));
    while (true) {
        bool want_ray = false;
        while (true) {
            // consume the last ray's result and advance to the next ray (or the pixel's end)
            while (state != kDone && !want_ray && !tracing) {
                switch (state) {
                case kNext:
                    if (c >= P.num_initial_candidates) {
                        bool stored = false;
                        if (Mc > 0 && have_sel) {
                            Reservoir r;
                            r.rc_pos = xyz(R.pos[i]); r.rc_n = xyz(R.nrm[i]); r.suffix = xyz(R.rad[i]);
                            r.prefix = mk(1.0f, 1.0f, 1.0f);
                            r.valid = 1.0f;
                            const float ph = target(pp, pn, load_mat(S, pmi), cp, r);
                            if (ph > 1e-10f) {
                                r.M = Mc;
                                r.w_sum = w_sum;
                                r.W = w_sum / ((float)Mc * ph);
                                store_res(R, i, r);
                                stored = true;
                            }
                        }
                        if (!stored) store_empty(R, i);
                        state = kDone;
                        break;
                    }
                    ++c;
                    dir = hash_primary_dir(G, pp, pn, pmi, cp, alpha, seed, src_pdf);
                    if (src_pdf <= 1e-6f || dot(pn, dir) <= 0.0f) break;  // dropped, not counted
                    ro = add(pp, scale(pn, bias));
                    rd = dir;
                    rtmax = 1e9f;
                    first_hit = false;
                    rays++;
                    state = kToRc;
                    want_ray = true;
                    break;
                case kToRc: {
                    if (h.prim == ~0u) { state = kNext; break; }  // escaped: dropped, not counted
                    const HitInfo rc = fetch_hit(S, h.prim, h.b1, h.b2);
                    rc_n = rc.normal;
                    if (dot(rc_n, dir) > 0.0f) rc_n = neg(rc_n);
                    rc_pos = rc.pos;
                    cur_mi = rc.mat;
                    const Mat m = load_mat(S, cur_mi);
                    suffix = rc_emission ? m.emissive : mk(0.0f, 0.0f, 0.0f);
                    cur_p = rc_pos; cur_n = rc_n;
                    thp = mk(1.0f, 1.0f, 1.0f);
                    b = 1;
                    if (light_sample(S, cur_p, cur_n, neg(dir), m, bias, seed, nee, ro, rd, rtmax)) {
                        first_hit = true;
                        shadow_rays++;
                        state = kShadow;
                        want_ray = true;
                    } else {
                        suffix = add(suffix, mul(thp, nee));
                        state = kBounce;
                    }
                    break;
                }
                case kShadow:
                    suffix = add(suffix, mul(thp, h.prim == ~0u ? nee : mk(0.0f, 0.0f, 0.0f)));
                    state = kBounce;
                    break;
                case kBounce: {
                    if (b >= max_b) { state = kFinish; break; }
                    f3 wi;
                    float pdf;
                    const float ux = rnd(seed), uy = rnd(seed);
                    const f3 f = mat_sample(load_mat(S, cur_mi), neg(dir), cur_n, ux, uy, wi, pdf);
                    if (pdf <= 1e-6f || dot(cur_n, wi) <= 0.0f) { state = kFinish; break; }
                    const float cosc = fmaxf(dot(cur_n, wi), 0.0f);
                    thp = mul(thp, divs(scale(f, cosc), pdf));
                    if (fmaxf(thp.x, fmaxf(thp.y, thp.z)) <= 0.0f) { state = kFinish; break; }
                    ro = add(cur_p, scale(cur_n, bias));
                    rd = wi;
                    dir = wi;
                    rtmax = 1e9f;
                    first_hit = false;
                    rays++;
                    ++b;
                    state = kExt;
                    want_ray = true;
                    break;
                }
                case kExt: {
                    if (h.prim == ~0u) {
                        suffix = add(suffix, mul(thp, sky(dir)));
                        state = kFinish;
                        break;
                    }
                    const HitInfo hi = fetch_hit(S, h.prim, h.b1, h.b2);
                    f3 hn = hi.normal;
                    if (dot(hn, dir) > 0.0f) hn = neg(hn);
                    cur_mi = hi.mat;
                    cur_p = hi.pos; cur_n = hn;
                    if (light_sample(S, cur_p, cur_n, neg(dir), load_mat(S, cur_mi), bias, seed, nee, ro, rd, rtmax)) {
                        first_hit = true;
                        shadow_rays++;
                        state = kShadow;
                        want_ray = true;
                    } else {
                        suffix = add(suffix, mul(thp, nee));
                        state = kBounce;
                    }
                    break;
                }
                case kFinish: {
                    Reservoir cand;
                    cand.rc_pos = rc_pos; cand.rc_n = rc_n; cand.suffix = suffix; cand.prefix = mk(1.0f, 1.0f, 1.0f);
                    cand.M = 1; cand.W = 0.0f; cand.w_sum = 0.0f; cand.valid = 1.0f;
                    const float ph = target(pp, pn, load_mat(S, pmi), cp, cand);
                    const f3 dl = sub(rc_pos, pp);
                    if (!(ph <= 0.0f) && !(dot(dl, dl) <= 1e-8f)) {
                        const float w = ph / fmaxf(src_pdf, 1e-6f);
                        w_sum += w;
                        Mc++;
                        if (w_sum > 0.0f && rnd(seed) < w / w_sum) {
                            R.pos[i] = make_float4(rc_pos.x, rc_pos.y, rc_pos.z, 0.0f);
                            R.nrm[i] = make_float4(rc_n.x, rc_n.y, rc_n.z, 0.0f);
                            R.rad[i] = make_float4(suffix.x, suffix.y, suffix.z, 0.0f);
                            have_sel = true;
                        }
                    }
                    state = kNext;
                    break;
                }
                default:
                    state = kDone;
                    break;
                }
            }
            // lanes without a pixel take the next one
            const bool need = state == kDone && !exhausted;
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
                            have_sel = false;
                            w_sum = 0.0f;
                            Mc = 0;
                            c = 0;
                            state = kNext;
                        } else {
                            hpos[i] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
                            hnrm[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                            store_empty(R, i);
                        }
                    }
                }
            }
        }
        if constexpr (steps) {
            if (want_ray) {
                tr = wray_setup(ro, rd);
                h.t = rtmax; h.prim = ~0u; h.b1 = h.b2 = 0.0f;
                tnode = 0;
                tsp = 0;
                tracing = true;
            }
            if (!__any(tracing)) break;  // every lane is out of pixels
            if (tracing && trace4q_steps<false, CULL>(S, tr, rd, bias, first_hit, h.t, h.prim, h.b1, h.b2, tnode, tsp, st,
                                                      step_visits))
                tracing = false;
            continue;
        } else {
            if (!__any(want_ray)) break;  // every lane is out of pixels
            if (want_ray) {
                if (S.speculative)
                    h = trace_closest4spec_s<false, true, CULL>(S, ro, rd, bias, rtmax, st, nullptr, nullptr, first_hit);
                else
                    h = trace_closest4_s<false, true, CULL>(S, ro, rd, bias, rtmax, st, nullptr, nullptr, first_hit);
            }
        }
    }
    count_rays(counter, rays);
    count_rays(counter + 1, shadow_rays);
}

// ---------------------------------------------------------------------------
// cs_temporal, restir_pt.slang:775-872 (reprojection :323-421)
// ---------------------------------------------------------------------------
template <int KS>
__global__ void __launch_bounds__(64) k_rpt_temporal(DevScene S, rrh_rpt_params P, Frame F, Res4 R, Res4 Rp,
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
        if (load_surface(F.gpos, F.gnrm, i, pp, pn, pmi)) {
            const bool found = find_history(cam, Hp, F.W, F.H, pp, pn, pmi, cp, hj, hp);
            if (found) {
                Reservoir cur = load_res(R, i);
                Reservoir prv = load_res(Rp, hj);
                if (prv.M != 0 && !(prv.valid < 0.5f)) {
                    uint32_t cap = 20u * P.num_initial_candidates;
                    cap = cap < 1u ? 1u : (cap > 8192u ? 8192u : cap);
                    prv.M = prv.M < cap ? prv.M : cap;
                    const float jac = jacobian(hp, pp, prv.rc_pos, prv.rc_n);
                    bool go = jac > 0.0f;
                    if (go) {
                        auto st = make_stack<STK>(s_stack);
                        rays++;
                        go = visible<CULL>(S, pp, pn, prv.rc_pos, P.ray_bias, st);
                    }
                    if (go) {
                        const Mat pm = load_mat(S, pmi);
                        const float ph = target(pp, pn, pm, cp, prv);
                        const float contrib = ph * fminf(prv.W, 50.0f) * (float)prv.M * jac;
                        if (contrib > 0.0f) {
                            cur.w_sum += contrib;
                            uint32_t seed = make_seed(px, py, P.frame_seed + 1u);
                            if (cur.w_sum > 0.0f && rnd(seed) < contrib / cur.w_sum) {
                                cur.rc_pos = prv.rc_pos; cur.rc_n = prv.rc_n;
                                cur.suffix = prv.suffix; cur.prefix = prv.prefix;
                                cur.valid = 1.0f;
                            }
                            const uint32_t nM = cur.M + prv.M;
                            if (nM > cap) {
                                cur.w_sum *= (float)cap / (float)nM;
                                cur.M = cap;
                            } else {
                                cur.M = nM;
                            }
                            if (cur.M > 0 && cur.valid > 0.5f) {
                                const float py_ = target(pp, pn, pm, cp, cur);
                                cur.W = (py_ > 1e-10f) ? (cur.w_sum / ((float)cur.M * py_)) : 0.0f;
                                cur.W = fminf(cur.W, 50.0f);
                            }
                            store_res(R, i, cur);
                        }
                    }
                }
            }
        }
    }
    count_rays(counter, rays);
}

// ---------------------------------------------------------------------------
// cs_spatial_shade, restir_pt.slang:876-1001
// ---------------------------------------------------------------------------
template <int KS>
__global__ void __launch_bounds__(64) k_rpt_spatial_shade(DevScene S, rrh_rpt_params P, HashGrid G, Frame F, Res4 R,
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
        const bool with_direct = P.include_direct_lighting != 0 && direct != nullptr;
        f3 L = mk(0.0f, 0.0f, 0.0f);
        if (with_direct) L = xyz(direct[i]);
        f3 pp, pn;
        uint32_t pmi;
        if (!load_surface(F.gpos, F.gnrm, i, pp, pn, pmi)) {
            if (P.include_direct_lighting == 0) L = add(L, sky(center_dir(cam, px, py)));
        } else {
            const Mat pm = load_mat(S, pmi);
            if (is_emissive(pm)) {
                L = add(L, pm.emissive);
            } else {
                Reservoir res = load_res(R, i);
                if (P.spatial_enable && P.spatial_num_neighbors > 0 && res.valid > 0.5f) {
                    const float phase = spiral_phase(px, py);
                    const float dist_scale = fmaxf(length(sub(pp, cp)), 1.0f);
                    for (uint32_t k = 0; k < P.spatial_num_neighbors; ++k) {
                        size_t j;
                        f3 q;
                        if (!neighbour(k, P.spatial_num_neighbors, P.spatial_radius, phase, px, py, F.W, F.H, j)) continue;
                        if (!neighbour_surface(F, j, pp, pn, dist_scale, q)) continue;
                        const Reservoir nb = load_res(R, j);
                        if (nb.M == 0 || nb.valid < 0.5f) continue;
                        const float jac = jacobian(q, pp, nb.rc_pos, nb.rc_n);
                        if (jac <= 0.0f) continue;
                        const float contrib = target(pp, pn, pm, cp, nb) * nb.W * (float)nb.M * jac;
                        if (contrib <= 0.0f) continue;
                        res.w_sum += contrib;
                        res.M += nb.M;
                        const float accept = fracf(phase + 0.754877666f * (float)(k + 1u));
                        if (res.w_sum > 0.0f && accept < contrib / res.w_sum) {
                            res.rc_pos = nb.rc_pos; res.rc_n = nb.rc_n;
                            res.suffix = nb.suffix; res.prefix = nb.prefix;
                        }
                    }
                    if (res.M > 0) {
                        const float ph = target(pp, pn, pm, cp, res);
                        res.W = (ph > 1e-10f) ? (res.w_sum / ((float)res.M * ph)) : 0.0f;
                    }
                }
                if (res.M > 0 && res.W > 0.0f && res.valid > 0.5f) {
                    const f3 ind = scale(contribution(pp, pn, pm, cp, res), res.W);
                    if (lum709(ind) > 0.0f) {
                        auto st = make_stack<STK>(s_stack);
                        rays++;
                        if (visible<CULL>(S, pp, pn, res.rc_pos, P.ray_bias, st)) L = add(L, scale(ind, P.indirect_strength));
                    }
                    // splat (restir_pt.slang:982-994): the direction toward the kept
                    // reconnection vertex trains next frame's guiding
                    if (P.guiding_enable && G.accum) {
                        const f3 to = sub(res.rc_pos, pp);
                        const float d2 = dot(to, to);
                        if (d2 > 1e-8f)
                            hash_splat(G.accum, hash_cell(pp, pmi, cp, G.base, G.slope, G.count),
                                       scale(to, 1.0f / sqrtf(d2)));
                    }
                }
            }
        }
        out[i] = make_float4(L.x, L.y, L.z, 1.0f);
    }
    count_rays(counter, rays);
}

__global__ void k_rpt_clear(Res4 R, float4* __restrict__ hpos, float4* __restrict__ hnrm, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    store_empty(R, i);
    hpos[i] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
    hnrm[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// counters: [0] extension rays of the initial stage, [1] its shadow rays and the temporal
// visibility rays, [2] the spatial stage's visibility rays (the primary surface comes from the
// G-buffer, so no primary rays), [3] the initial stage's next pixel
__global__ void k_rpt_totals(const uint32_t* __restrict__ cnt, unsigned long long* __restrict__ tot) {
    if (threadIdx.x != 0) return;
    tot[1] += cnt[0];
    tot[2] += cnt[1] + cnt[2];
}

// ---- host side -----------------------------------------------------------------------
int res4(rrh_image* const* r, uint32_t W, uint32_t H, Res4& out, const char* who) {
    if (!r) { set_error(std::string(who) + ": null reservoir"); return RRH_E_ARG; }
    for (int k = 0; k < 4; ++k)
        if (!fits(r[k], W, H)) {
            set_error(std::string(who) + ": reservoir planes must be four RGBA32F images of the camera's extent");
            return RRH_E_ARG;
        }
    out = Res4{(float4*)r[0]->p, (float4*)r[1]->p, (float4*)r[2]->p, (float4*)r[3]->p};
    return RRH_OK;
}

int hash_grid(rrh_grid* g, HashGrid& G, const char* who) {
    G = HashGrid{};
    if (!g) return RRH_OK;
    if (g->p.kind != RRH_GRID_HASH) { set_error(std::string(who) + ": guiding grid must be of kind RRH_GRID_HASH"); return RRH_E_ARG; }
    G.fitted = (const float4*)g->fitted.p;
    G.accum = (uint32_t*)g->accum.p;
    G.base = g->p.base_cell_size;
    G.slope = g->p.distance_slope;
    G.count = g->p.cell_count;
    return RRH_OK;
}

int prepare(rrh_ctx* c, rrh_scene* s, const rrh_rpt_params* p, const rrh_image* gpos, const rrh_image* gnrm,
            Frame& F, const char* who) {
    if (!p) { set_error(std::string(who) + ": null argument"); return RRH_E_ARG; }
    int rc = check_frame(c, s, gpos, gnrm, p->num_initial_candidates, p->spatial_num_neighbors, F, who);
    if (rc) return rc;
    if (p->max_bounces > 64) { set_error(std::string(who) + ": max_bounces <= 64"); return RRH_E_RANGE; }
    return c->rpt_counter.ensure(16);
}

template <int S> struct KInit { static constexpr auto fn = k_rpt_initial<S, false>; };
template <int S> struct KInitSteps { static constexpr auto fn = k_rpt_initial<S, true>; };
template <int S> struct KTemp { static constexpr auto fn = k_rpt_temporal<S>; };
template <int S> struct KSpat { static constexpr auto fn = k_rpt_spatial_shade<S>; };

}  // namespace

extern "C" {

int rrh_rpt_initial(rrh_ctx* c, rrh_scene* s, const rrh_rpt_params* p, rrh_grid* grid, const rrh_image* gpos,
                    const rrh_image* gnrm, rrh_image* const res[4], rrh_image* hist_pos, rrh_image* hist_norm) {
    Frame F;
    int rc = prepare(c, s, p, gpos, gnrm, F, "rrh_rpt_initial");
    if (rc) return rc;
    RRH_ARG(fits(hist_pos, F.W, F.H) && fits(hist_norm, F.W, F.H),
            "rrh_rpt_initial: history position / normal must be RGBA32F images of the camera's extent");
    Res4 R;
    if ((rc = res4(res, F.W, F.H, R, "rrh_rpt_initial"))) return rc;
    HashGrid G;
    if ((rc = hash_grid(grid, G, "rrh_rpt_initial"))) return rc;
    RRH_HIP(hipMemsetAsync(c->rpt_counter.p, 0, 16, c->stream));
    // persistent waves: as many as the CUs hold at once (the occupancy the kernel's registers and
    // LDS stack allow), never more than there are 8x8 tiles
    const int sk = kernel_kind(c, s);
    const uint32_t tiles_x = (F.W + 7) / 8, tiles = tiles_x * ((F.H + 7) / 8);
    uint32_t waves = 0;
    if (c->rpt_step_visits > 0 && c->speculative) {  // traversal in passes (RRH_OPT_RPT_STEP_VISITS)
        rc = persistent_waves<KInitSteps>(c, sk, tiles, waves);
        if (rc) return rc;
        launch<KInitSteps>(sk, dim3(waves), c->stream, s->dev(), *p, G, F, R, (float4*)hist_pos->p,
                           (float4*)hist_norm->p, (uint32_t*)c->rpt_counter.p, tiles_x, tiles * 64u, c->rpt_step_visits);
    } else {
        rc = persistent_waves<KInit>(c, sk, tiles, waves);
        if (rc) return rc;
        launch<KInit>(sk, dim3(waves), c->stream, s->dev(), *p, G, F, R, (float4*)hist_pos->p, (float4*)hist_norm->p,
                      (uint32_t*)c->rpt_counter.p, tiles_x, tiles * 64u, 0u);
    }
    RRH_LAUNCH_CHECK();
    return RRH_OK;
}

int rrh_rpt_temporal(rrh_ctx* c, rrh_scene* s, const rrh_rpt_params* p, const rrh_image* gpos, const rrh_image* gnrm,
                     rrh_image* const res[4], rrh_image* const res_prev[4], const rrh_image* hist_prev_pos,
                     const rrh_image* hist_prev_norm) {
    Frame F;
    int rc = prepare(c, s, p, gpos, gnrm, F, "rrh_rpt_temporal");
    if (rc) return rc;
    RRH_ARG(fits(hist_prev_pos, F.W, F.H) && fits(hist_prev_norm, F.W, F.H),
            "rrh_rpt_temporal: history position / normal must be RGBA32F images of the camera's extent");
    Res4 R, Rp;
    if ((rc = res4(res, F.W, F.H, R, "rrh_rpt_temporal"))) return rc;
    if ((rc = res4(res_prev, F.W, F.H, Rp, "rrh_rpt_temporal"))) return rc;
    if (!p->temporal_enable) return RRH_OK;
    launch<KTemp>(kernel_kind(c, s), dim3((F.W + 7) / 8, (F.H + 7) / 8), c->stream, s->dev(), *p, F, R, Rp,
                  History{(const float4*)hist_prev_pos->p, (const float4*)hist_prev_norm->p},
                  (uint32_t*)c->rpt_counter.p + 1, c->xcd_remap);
    RRH_LAUNCH_CHECK();
    return RRH_OK;
}

int rrh_rpt_spatial_shade(rrh_ctx* c, rrh_scene* s, const rrh_rpt_params* p, rrh_grid* grid, const rrh_image* gpos,
                          const rrh_image* gnrm, const rrh_image* direct, rrh_image* const res[4], rrh_image* out) {
    Frame F;
    int rc = prepare(c, s, p, gpos, gnrm, F, "rrh_rpt_spatial_shade");
    if (rc) return rc;
    RRH_ARG(fits(out, F.W, F.H), "rrh_rpt_spatial_shade: output must be an RGBA32F image of the camera's extent");
    RRH_ARG(!direct || fits(direct, F.W, F.H),
            "rrh_rpt_spatial_shade: direct lighting must be an RGBA32F image of the camera's extent");
    Res4 R;
    if ((rc = res4(res, F.W, F.H, R, "rrh_rpt_spatial_shade"))) return rc;
    HashGrid G;
    if ((rc = hash_grid(grid, G, "rrh_rpt_spatial_shade"))) return rc;
    launch<KSpat>(kernel_kind(c, s), dim3((F.W + 7) / 8, (F.H + 7) / 8), c->stream, s->dev(), *p, G, F, R,
                  direct ? (const float4*)direct->p : (const float4*)nullptr, (float4*)out->p,
                  (uint32_t*)c->rpt_counter.p + 2, c->xcd_remap);
    RRH_LAUNCH_CHECK();
    if (!c->totals.p) {
        if ((rc = c->totals.ensure(32))) return rc;
        RRH_HIP(hipMemsetAsync(c->totals.p, 0, 32, c->stream));
    }
    hipLaunchKernelGGL(k_rpt_totals, dim3(1), dim3(64), 0, c->stream, (const uint32_t*)c->rpt_counter.p,
                       (unsigned long long*)c->totals.p);
    RRH_LAUNCH_CHECK();
    return RRH_OK;
}

int rrh_rpt_clear(rrh_ctx* c, rrh_image* const res[4], rrh_image* hist_pos, rrh_image* hist_norm) {
    RRH_ARG(c && res && hist_pos && hist_norm, "rrh_rpt_clear: null argument");
    const uint32_t W = hist_pos->w, H = hist_pos->h;
    RRH_ARG(fits(hist_pos, W, H) && fits(hist_norm, W, H), "rrh_rpt_clear: history planes must be RGBA32F, same extent");
    Res4 R;
    int rc = res4(res, W, H, R, "rrh_rpt_clear");
    if (rc) return rc;
    const uint32_t n = W * H;
    if (n == 0) return RRH_OK;
    hipLaunchKernelGGL(k_rpt_clear, dim3((n + 255) / 256), dim3(256), 0, c->stream, R, (float4*)hist_pos->p,
                       (float4*)hist_norm->p, n);
    RRH_LAUNCH_CHECK();
    return RRH_OK;
}

int rrh_rpt_ray_counts(rrh_ctx* c, rrh_ray_counts* out) {
    RRH_ARG(c && out, "rrh_rpt_ray_counts: null argument");
    out->primary = out->extension = out->shadow = 0;
    if (!c->rpt_counter.p) return RRH_OK;
    uint32_t h[4] = {0, 0, 0, 0};
    RRH_HIP(hipMemcpyAsync(h, c->rpt_counter.p, 16, hipMemcpyDeviceToHost, c->stream));
    RRH_HIP(hipStreamSynchronize(c->stream));
    out->extension = h[0];
    out->shadow = (uint64_t)h[1] + h[2];
    return RRH_OK;
}

}  // extern "C"

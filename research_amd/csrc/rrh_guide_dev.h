// rrh_guide_dev.h -- path-guiding pieces used inside the path tracer's shade kernel:
// one-sample MIS between the BSDF and a vMF mixture (8-lobe EM grid or the learned
// guide), and the path-vertex records that train both guides.
#pragma once
#include <hip/hip_fp16.h>

#include "rrh_device.h"

struct rrh_ctx;
struct rrh_scene;

namespace rrh {

constexpr uint32_t kMaxLobes = 8;

// ---- the reference's single-lobe hash grid (GuidingGrid.slang), shared by the grid's own
// kernels and ReSTIR PT ---------------------------------------------------------------------
struct HashGrid {
    const float4* fitted = nullptr;  // per cell (mu, kappa); nullptr: no guiding
    uint32_t* accum = nullptr;       // per cell (sum x, y, z in 2^-20 units, count)
    float base = 0.0f, slope = 0.0f;
    uint32_t count = 1;
};

// GuidingGrid.slang:54-67: cell edge grows with the distance to the camera; the quantised
// position and the BSDF id hash to one of `count` cells
RRH_D uint32_t hash_cell(f3 p, uint32_t bsdf, f3 cam, float base, float slope, uint32_t count) {
    float vd = length(sub(p, cam));
    float cell = fmaxf(base + slope * vd, 1.0e-4f);
    int qx = (int)floorf(p.x / cell), qy = (int)floorf(p.y / cell), qz = (int)floorf(p.z / cell);
    uint32_t h = pcg_hash((uint32_t)qx * 73856093u ^ (uint32_t)qy * 19349663u ^ (uint32_t)qz * 83492791u ^
                          bsdf * 2654435761u);
    return h % (count > 1u ? count : 1u);
}

// GuidingGrid.slang:73-90: each component rounded to the nearest 2^-20 and added as a two's
// complement integer (order-independent sums), plus the sample count
RRH_D void hash_splat(uint32_t* accum, uint32_t cell, f3 dir) {
    uint32_t* a = accum + (size_t)cell * 4;
    atomicAdd(a + 0, (uint32_t)(int32_t)roundf(dir.x * 1048576.0f));
    atomicAdd(a + 1, (uint32_t)(int32_t)roundf(dir.y * 1048576.0f));
    atomicAdd(a + 2, (uint32_t)(int32_t)roundf(dir.z * 1048576.0f));
    atomicAdd(a + 3, 1u);
}

// sample_primary_dir, restir_pt.slang:520-563: one-sample MIS between the cosine hemisphere and
// the cell's fitted lobe (GuidingGrid.slang:147-157 decides whether the lobe is usable);
// src_pdf = alpha p_vmf + (1 - alpha) p_cos at the chosen direction
RRH_D f3 hash_primary_dir(const HashGrid& G, f3 p, f3 N, uint32_t bsdf, f3 cam, float alpha_in, uint32_t& seed,
                          float& src_pdf) {
    float alpha = 0.0f;
    bool valid = false;
    f3 mu = mk(0.0f, 0.0f, 0.0f);
    float kappa = 0.0f;
    if (G.fitted && alpha_in > 0.0f) {
        const float4 f = G.fitted[hash_cell(p, bsdf, cam, G.base, G.slope, G.count)];
        mu = xyz(f);
        kappa = f.w;
        valid = (kappa > 0.0f) && (dot(mu, mu) > 0.25f);
        if (valid) alpha = alpha_in;
    }
    f3 T, B;
    build_onb(N, T, B);
    f3 wi;
    if (rnd(seed) < alpha) {
        float u0 = rnd(seed), u1 = rnd(seed), vp;
        wi = vmf_sample(mu, kappa, u0, u1, vp);
    } else {
        float u0 = rnd(seed), u1 = rnd(seed), cp;
        wi = to_world(cosine_hemisphere(u0, u1, cp), T, B, N);
    }
    float ci = dot(N, wi);
    float pc = (ci > 0.0f) ? (ci * K_INV_PI) : 0.0f;
    float pv = valid ? vmf_pdf(mu, kappa, wi) : 0.0f;
    src_pdf = alpha * pv + (1.0f - alpha) * pc;
    return wi;
}

struct GuideDev {
    const float* lobes = nullptr;         // VMM: res^3 cells x K x (mu3, kappa, pi)
    const uint16_t* mlp_lobes = nullptr;  // MLP: per ray-queue slot kMlpLobeHalves fp16 (LobesH16)
    uint32_t res = 0, K = 0;
    float bmin[3] = {0, 0, 0}, bext[3] = {1, 1, 1};
};

// path-vertex record (8 float4 = 128 B) at index (pixel >> shift) * max_bounces + bounce, for the
// pixels with pixel mod 2^shift = phase (whole paths of a pixel lattice; the phase moves every
// recorded frame):
//  r0 (pos, stamp) r1 (n, pdf) r2 (wo, mat) r3 (wi, nee visible)
//  r4 (L before this vertex, alive) r5 (NEE contribution if visible) r6 (throughput after)
//  r7 (BSDF weight f cos / pdf before Russian roulette)
struct VertexRec {
    float4* v = nullptr;
    uint32_t max_bounces = 0;
    uint32_t stamp = 0;
    uint32_t shift = 0, phase = 0;
};
constexpr uint32_t kRecF4 = 8;

RRH_D uint32_t vmm_cell(const GuideDev& G, f3 p) {
    float f[3] = {(p.x - G.bmin[0]) / G.bext[0] * (float)G.res, (p.y - G.bmin[1]) / G.bext[1] * (float)G.res,
                  (p.z - G.bmin[2]) / G.bext[2] * (float)G.res};
    uint32_t c[3];
    for (int k = 0; k < 3; ++k) {
        int q = (int)floorf(f[k]);
        q = q < 0 ? 0 : (q > (int)G.res - 1 ? (int)G.res - 1 : q);
        c[k] = (uint32_t)q;
    }
    return (c[2] * G.res + c[1]) * G.res + c[0];
}

// lobe accessors of the one-sample MIS below: fp32 (mu3, kappa, pi) records (the EM grid's cell),
// or the learned guide's fp16 record of one queue slot: 8 x (mu.xyz, kappa) then the 8 weights.
// A fp16 mu is renormalised wherever it is read, so sampling and the pdf use the same unit
// directions and the estimate stays unbiased with the rounded lobes.
struct LobesF32 {
    const float* p;
    RRH_D float pi(uint32_t k) const { return p[k * 5 + 4]; }
    RRH_D f3 mu(uint32_t k) const { return mk(p[k * 5], p[k * 5 + 1], p[k * 5 + 2]); }
    RRH_D float kappa(uint32_t k) const { return p[k * 5 + 3]; }
};
constexpr uint32_t kMlpLobeHalves = 40;
RRH_D float half_bits_to_float(uint32_t h) { return __half2float(__ushort_as_half((unsigned short)h)); }
struct LobesH16 {
    const uint32_t* p;  // 20 words: (mu.x, mu.y), (mu.z, kappa) per lobe, then the weights in pairs
    RRH_D float pi(uint32_t k) const {
        const uint32_t w = p[16 + (k >> 1)];
        return half_bits_to_float((k & 1u) ? (w >> 16) : (w & 0xffffu));
    }
    RRH_D f3 mu(uint32_t k) const {
        const uint32_t a = p[2 * k], b = p[2 * k + 1];
        return normalize(mk(half_bits_to_float(a & 0xffffu), half_bits_to_float(a >> 16), half_bits_to_float(b & 0xffffu)));
    }
    RRH_D float kappa(uint32_t k) const { return half_bits_to_float(p[2 * k + 1] >> 16); }
};

template <class Lobes>
RRH_D float mixture_pdf(const Lobes& lc, uint32_t K, f3 wi) {
    float p = 0.0f;
    for (uint32_t k = 0; k < K; ++k) {
        float pk = lc.pi(k);
        if (pk > 0.0f) p += pk * vmf_pdf(lc.mu(k), lc.kappa(k), wi);
    }
    return p;
}

// one-sample MIS with the mixture lc (K lobes; K = 0: no guide): with probability alpha sample
// the mixture, else the BSDF; pdf = alpha * p_guide + (1 - alpha) * p_bsdf; returns f(wo, wi)
template <class Lobes>
RRH_D f3 mixture_guided_sample(const Lobes& lc, uint32_t K, float alpha, f3 n, f3 wo, const Mat& m, uint32_t& seed,
                               f3& wi, float& pdf) {
    float tot = 0.0f;
    for (uint32_t k = 0; k < K; ++k) tot += fmaxf(lc.pi(k), 0.0f);
    float a = (tot > 0.0f) ? alpha : 0.0f;
    float usel = rnd(seed);
    if (usel < a) {
        float uk = rnd(seed) * tot;
        uint32_t pick = K - 1;
        float acc = 0.0f;
        for (uint32_t k = 0; k < K; ++k) {
            acc += fmaxf(lc.pi(k), 0.0f);
            if (uk < acc) { pick = k; break; }
        }
        float u0 = rnd(seed), u1 = rnd(seed), vp;
        wi = vmf_sample(lc.mu(pick), lc.kappa(pick), u0, u1, vp);
    } else {
        float ux = rnd(seed), uy = rnd(seed), bp;
        f3 f = mat_sample(m, wo, n, ux, uy, wi, bp);
        if (!m.lambert && bp == 1.0f && f.x == 0.0f && f.y == 0.0f && f.z == 0.0f) { pdf = 0.0f; return f; }
    }
    float pb = mat_sample_pdf(m, wo, wi, n);
    float pg = (a > 0.0f) ? mixture_pdf(lc, K, wi) / tot : 0.0f;
    pdf = a * pg + (1.0f - a) * pb;
    if (dot(n, wi) <= 0.0f) { pdf = 0.0f; return mk(0, 0, 0); }
    return mat_eval(m, n, wo, wi);
}

// slot: the ray's queue index (the learned guide's lobes are inferred per queue slot)
RRH_D f3 guided_sample(const GuideDev& G, uint32_t mode, float alpha, const DevScene& S, uint32_t slot, f3 pos, f3 n,
                       f3 wo, const Mat& m, uint32_t& seed, f3& wi, float& pdf) {
    (void)S;
    if (mode == RRH_GUIDE_MLP && G.mlp_lobes)
        return mixture_guided_sample(LobesH16{reinterpret_cast<const uint32_t*>(G.mlp_lobes + (size_t)slot * kMlpLobeHalves)},
                                     kMaxLobes, alpha, n, wo, m, seed, wi, pdf);
    const bool vmm = mode == RRH_GUIDE_VMM && G.lobes;
    return mixture_guided_sample(LobesF32{vmm ? G.lobes + (size_t)vmm_cell(G, pos) * G.K * 5 : nullptr}, vmm ? G.K : 0u,
                                 alpha, n, wo, m, seed, wi, pdf);
}

RRH_D bool guide_records(const VertexRec& R, uint32_t pixel) { return (pixel & ((1u << R.shift) - 1u)) == R.phase; }
RRH_D void guide_record_vertex(const VertexRec& R, uint32_t pixel, uint32_t b, f3 pos, f3 n, f3 wo, f3 wi, float pdf,
                               f3 bsdf_w, f3 thp_after, f3 L_pre, f3 nee_c, uint32_t mat, bool alive) {
    if (!guide_records(R, pixel)) return;
    float4* r = R.v + ((size_t)(pixel >> R.shift) * R.max_bounces + b) * kRecF4;
    r[0] = make_float4(pos.x, pos.y, pos.z, __uint_as_float(R.stamp));
    r[1] = make_float4(n.x, n.y, n.z, pdf);
    r[2] = make_float4(wo.x, wo.y, wo.z, __uint_as_float(mat));
    r[3] = make_float4(wi.x, wi.y, wi.z, 0.0f);
    r[4] = make_float4(L_pre.x, L_pre.y, L_pre.z, alive ? 1.0f : 0.0f);
    r[5] = make_float4(nee_c.x, nee_c.y, nee_c.z, 0.0f);
    r[6] = make_float4(thp_after.x, thp_after.y, thp_after.z, 0.0f);
    r[7] = make_float4(bsdf_w.x, bsdf_w.y, bsdf_w.z, 0.0f);
}
RRH_D void guide_record_nee_visible(const VertexRec& R, uint32_t pixel, uint32_t b) {
    if (!guide_records(R, pixel)) return;
    float4* r = R.v + ((size_t)(pixel >> R.shift) * R.max_bounces + b) * kRecF4;
    r[3].w = 1.0f;
}

// The learned guide's input record of queue slot i: the hit point (w = 1), the normal facing
// the ray origin with the material index, and the outgoing direction; zeros for a miss.  d is
// the slot's ray direction, h its hit (t, prim bits, b1, b2).  The input pass (k_mlp_prep) and
// the traversal kernels' epilogue both write it through this function, so their records agree
// bit for bit.
RRH_D void write_hit_prep(const DevScene& S, float4 d, float4 h, float4* __restrict__ prep, size_t i) {
    const uint32_t prim = __float_as_uint(h.y);
    float4 p0 = make_float4(0, 0, 0, 0), p1 = p0, p2 = p0;
    if (prim != ~0u) {
        HitInfo hi = fetch_hit(S, prim, h.z, h.w);
        f3 wo = neg(xyz(d));
        f3 nn = hi.normal;
        if (dot(nn, wo) < 0.0f) nn = neg(nn);
        p0 = make_float4(hi.pos.x, hi.pos.y, hi.pos.z, 1.0f);
        p1 = make_float4(nn.x, nn.y, nn.z, __uint_as_float(hi.mat));
        p2 = make_float4(wo.x, wo.y, wo.z, 0.0f);
    }
    prep[i * 3] = p0;
    prep[i * 3 + 1] = p1;
    prep[i * 3 + 2] = p2;
}

// host side (rrh_guide.hip)
int guide_prepare(rrh_ctx* c, rrh_scene* s, const rrh_pt_params* p, rrh_grid* grid, rrh_mlp* mlp, GuideDev& G,
                  VertexRec& R);
int guide_finish(rrh_ctx* c, rrh_scene* s, const rrh_pt_params* p, rrh_grid* grid, rrh_mlp* mlp, const VertexRec& R,
                 const float4* rad);
// learned guide: the fp16 lobe record (kMlpLobeHalves) of every hit of queue b (rrh_mlp.hip)
// (prep_done: the traversal launch already wrote the input records into c->mlp_prep)
int mlp_infer_queue(rrh_ctx* c, rrh_mlp* m, const DevScene& S, const float4* qo, const float4* qd, const float4* hits,
                    const uint32_t* cnt, uint32_t b, uint32_t cap, uint16_t* lobes, bool prep_done = false);

}  // namespace rrh

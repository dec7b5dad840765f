// rrh_atrous.hip -- edge-avoiding A-trous wavelet filter.
//
// Reference mode restates atrous.slang cs_main :41-106 (Dammertz et al. 2010 weights:
// Gaussian on |dpos|^2, 1 - n.n and dluma^2 with Rec.601 luma, B3 5x5 taps, clamp to
// edge, miss taps skipped) driven like AtrousPass::execute (src/passes/denoise/
// AtrousPass.cpp:243-306: iterations clamped to [1,7], stride 1<<i, ping-pong, the
// result in slot (iterations-1)&1).
// Config mode is the SVGF edge-stopping of Schied et al. 2017 (eqs. 3-5) on fp16
// buffers: luminance term scaled by the 3x3-Gaussian-filtered variance carried in
// alpha, normal term max(0, n.n)^sigma_n, depth term against the screen-space
// depth gradient; the variance is filtered with squared weights.
//
// Both are HBM streams: each output pixel reads 25 taps whose footprint is shared
// with its neighbours, so a 2D tile of outputs per workgroup keeps the taps in L2;
// for strides 1 and 2 the tile plus halo is staged in LDS first.
#include <hip/hip_fp16.h>

#include "rrh_internal.h"

using namespace rrh;

namespace {

constexpr int TX = 16, TY = 16;

__device__ __forceinline__ float kw(int o) { return (o == -2 || o == 2) ? 1.0f : ((o == -1 || o == 1) ? 4.0f : 6.0f); }
__device__ __forceinline__ float luma601(f3 c) { return c.x * 0.299f + c.y * 0.587f + c.z * 0.114f; }
__device__ __forceinline__ int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

struct RefPx { f3 pos; float w; f3 nrm; f3 rgb; };

// ---- reference mode -----------------------------------------------------------------
// LDS-staged variant: the (TX + 4s) x (TY + 4s) window of position/normal/radiance
// around the tile is loaded once (clamped to the image), then all 25 taps read LDS.
template <int S>
__global__ void __launch_bounds__(TX* TY) k_atrous_ref_lds(int W, int H, float sps, float sns, float sls,
                                                          const float4* __restrict__ in, const float4* __restrict__ gpos,
                                                          const float4* __restrict__ gnrm, float4* __restrict__ out) {
    constexpr int HW = TX + 4 * S, HH = TY + 4 * S;
    __shared__ float4 s_pos[HH * HW];
    __shared__ float4 s_nrm[HH * HW];
    __shared__ float4 s_rad[HH * HW];
    const int x0 = blockIdx.x * TX - 2 * S, y0 = blockIdx.y * TY - 2 * S;
    for (int k = threadIdx.x; k < HW * HH; k += TX * TY) {
        int lx = k % HW, ly = k / HW;
        int gx = clampi(x0 + lx, 0, W - 1), gy = clampi(y0 + ly, 0, H - 1);
        size_t g = (size_t)gy * W + gx;
        s_pos[k] = gpos[g];
        s_nrm[k] = gnrm[g];
        s_rad[k] = in[g];
    }
    __syncthreads();
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    const int x = blockIdx.x * TX + tx, y = blockIdx.y * TY + ty;
    if (x >= W || y >= H) return;
    const int lc = (ty + 2 * S) * HW + (tx + 2 * S);
    float4 cp = s_pos[lc], cr = s_rad[lc];
    size_t i = (size_t)y * W + x;
    if (cp.w < 0.0f) { out[i] = cr; return; }
    f3 cn = normalize(xyz(s_nrm[lc]));
    float cl = luma601(xyz(cr));
    f3 acc = mk(0, 0, 0);
    float wsum = 0.0f;
    for (int ty2 = -2; ty2 <= 2; ++ty2) {
        for (int tx2 = -2; tx2 <= 2; ++tx2) {
            // clamp in image space, then map into the staged window
            int sx = clampi(x + tx2 * S, 0, W - 1), sy = clampi(y + ty2 * S, 0, H - 1);
            int l = (sy - y0) * HW + (sx - x0);
            float4 sp = s_pos[l];
            if (sp.w < 0.0f) continue;
            f3 sr = xyz(s_rad[l]);
            float bw = (kw(tx2) * kw(ty2)) / 256.0f;
            f3 pd = sub(xyz(sp), xyz(cp));
            float pw = exp_rr(-dot(pd, pd) / sps);
            f3 sn = normalize(xyz(s_nrm[l]));
            float nd = fmaxf(0.0f, 1.0f - dot(cn, sn));
            float nw = exp_rr(-nd / sns);
            float ld = luma601(sr) - cl;
            float lw = exp_rr(-(ld * ld) / sls);
            float w = bw * pw * nw * lw;
            acc = add(acc, scale(sr, w));
            wsum += w;
        }
    }
    f3 r = (wsum > 0.0f) ? divs(acc, wsum) : xyz(cr);
    out[i] = make_float4(r.x, r.y, r.z, cr.w);
}

__global__ void __launch_bounds__(TX* TY) k_atrous_ref(int W, int H, int st, float sps, float sns, float sls,
                                                      const float4* __restrict__ in, const float4* __restrict__ gpos,
                                                      const float4* __restrict__ gnrm, float4* __restrict__ out) {
    const int x = blockIdx.x * TX + (int)(threadIdx.x % TX), y = blockIdx.y * TY + (int)(threadIdx.x / TX);
    if (x >= W || y >= H) return;
    size_t i = (size_t)y * W + x;
    float4 cp = gpos[i], cr = in[i];
    if (cp.w < 0.0f) { out[i] = cr; return; }
    f3 cn = normalize(xyz(gnrm[i]));
    float cl = luma601(xyz(cr));
    f3 acc = mk(0, 0, 0);
    float wsum = 0.0f;
    for (int ty = -2; ty <= 2; ++ty) {
        for (int tx = -2; tx <= 2; ++tx) {
            int sx = clampi(x + tx * st, 0, W - 1), sy = clampi(y + ty * st, 0, H - 1);
            size_t j = (size_t)sy * W + sx;
            float4 sp = gpos[j];
            if (sp.w < 0.0f) continue;
            f3 sr = xyz(in[j]);
            float bw = (kw(tx) * kw(ty)) / 256.0f;
            f3 pd = sub(xyz(sp), xyz(cp));
            float pw = exp_rr(-dot(pd, pd) / sps);
            f3 sn = normalize(xyz(gnrm[j]));
            float nd = fmaxf(0.0f, 1.0f - dot(cn, sn));
            float nw = exp_rr(-nd / sns);
            float ld = luma601(sr) - cl;
            float lw = exp_rr(-(ld * ld) / sls);
            float w = bw * pw * nw * lw;
            acc = add(acc, scale(sr, w));
            wsum += w;
        }
    }
    f3 r = (wsum > 0.0f) ? divs(acc, wsum) : xyz(cr);
    out[i] = make_float4(r.x, r.y, r.z, cr.w);
}

// ---- config mode (fp16) ---------------------------------------------------------------
__device__ __forceinline__ float h2f(unsigned short h) { return __half2float(__ushort_as_half(h)); }
__device__ __forceinline__ unsigned short f2h(float f) { return __half_as_ushort(__float2half_rn(f)); }

struct H4 { unsigned short x, y, z, w; };
__device__ __forceinline__ H4 ldh(const ushort4* p, size_t i) { ushort4 v = p[i]; return H4{v.x, v.y, v.z, v.w}; }
__device__ __forceinline__ float powpos(float x, float e) { return (x > 0.0f) ? exp_rr(e * log_rr(x)) : 0.0f; }

__global__ void __launch_bounds__(TX* TY) k_atrous_cfg(int W, int H, int st, float sz, float sn, float sl,
                                                      const ushort4* __restrict__ in, const ushort4* __restrict__ gb,
                                                      ushort4* __restrict__ out) {
    const int x = blockIdx.x * TX + (int)(threadIdx.x % TX), y = blockIdx.y * TY + (int)(threadIdx.x / TX);
    if (x >= W || y >= H) return;
    size_t i = (size_t)y * W + x;
    H4 cg = ldh(gb, i);
    float zp = h2f(cg.w);
    if (!(zp > 0.0f)) { out[i] = in[i]; return; }
    f3 np = mk(h2f(cg.x), h2f(cg.y), h2f(cg.z));
    H4 ci = ldh(in, i);
    float lp = lum709(mk(h2f(ci.x), h2f(ci.y), h2f(ci.z)));
    const float gk[3] = {0.25f, 0.5f, 0.25f};
    float var = 0.0f;
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            int sx = clampi(x + dx, 0, W - 1), sy = clampi(y + dy, 0, H - 1);
            var += gk[dx + 1] * gk[dy + 1] * h2f(in[(size_t)sy * W + sx].w);
        }
    float phil = sl * sqrtf(fmaxf(0.0f, var)) + 1e-10f;
    int xm = x > 0 ? x - 1 : 0, xp = x < W - 1 ? x + 1 : W - 1;
    int ym = y > 0 ? y - 1 : 0, yp = y < H - 1 ? y + 1 : H - 1;
    float dzdx = (h2f(gb[(size_t)y * W + xp].w) - h2f(gb[(size_t)y * W + xm].w)) * 0.5f;
    float dzdy = (h2f(gb[(size_t)yp * W + x].w) - h2f(gb[(size_t)ym * W + x].w)) * 0.5f;
    f3 acc = mk(0, 0, 0);
    float accv = 0.0f, wsum = 0.0f;
    for (int ty = -2; ty <= 2; ++ty) {
        for (int tx = -2; tx <= 2; ++tx) {
            int sx = clampi(x + tx * st, 0, W - 1), sy = clampi(y + ty * st, 0, H - 1);
            size_t j = (size_t)sy * W + sx;
            H4 sg = ldh(gb, j);
            float zq = h2f(sg.w);
            if (!(zq > 0.0f)) continue;
            f3 nq = mk(h2f(sg.x), h2f(sg.y), h2f(sg.z));
            H4 si = ldh(in, j);
            f3 cq = mk(h2f(si.x), h2f(si.y), h2f(si.z));
            float vq = h2f(si.w);
            float bw = (kw(tx) * kw(ty)) / 256.0f;
            float wz = fabsf(zp - zq) / (sz * fabsf(dzdx * (float)(tx * st) + dzdy * (float)(ty * st)) + 1e-3f);
            float wn = powpos(fmaxf(0.0f, dot(np, nq)), sn);
            float wl = fabsf(lp - lum709(cq)) / phil;
            float w = bw * exp_rr(0.0f - wl - wz) * wn;
            acc = add(acc, scale(cq, w));
            accv += vq * (w * w);
            wsum += w;
        }
    }
    if (wsum > 0.0f) {
        ushort4 o;
        o.x = f2h(acc.x / wsum); o.y = f2h(acc.y / wsum); o.z = f2h(acc.z / wsum); o.w = f2h(accv / (wsum * wsum));
        out[i] = o;
    } else {
        out[i] = in[i];
    }
}

// config-mode inputs: rgb + 3x3 luminance variance (fp16), normal + linear depth (fp16)
__global__ void __launch_bounds__(TX* TY) k_atrous_prepare(int W, int H, f3 cpos, f3 cdir, const float4* __restrict__ rad,
                                                          const float4* __restrict__ gpos, const float4* __restrict__ gnrm,
                                                          ushort4* __restrict__ rad16, ushort4* __restrict__ gb16) {
    const int x = blockIdx.x * TX + (int)(threadIdx.x % TX), y = blockIdx.y * TY + (int)(threadIdx.x / TX);
    if (x >= W || y >= H) return;
    size_t i = (size_t)y * W + x;
    float4 c = rad[i];
    float m1 = 0.0f, m2 = 0.0f, n = 0.0f;
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            int sx = clampi(x + dx, 0, W - 1), sy = clampi(y + dy, 0, H - 1);
            size_t j = (size_t)sy * W + sx;
            if (gpos[j].w < 0.0f) continue;
            float l = lum709(xyz(rad[j]));
            m1 += l; m2 += l * l; n += 1.0f;
        }
    float var = 0.0f;
    if (n > 0.0f) { m1 = m1 / n; m2 = m2 / n; var = fmaxf(0.0f, m2 - m1 * m1); }
    ushort4 r;
    r.x = f2h(c.x); r.y = f2h(c.y); r.z = f2h(c.z); r.w = f2h(var);
    rad16[i] = r;
    float4 gp = gpos[i];
    float4 gn = gnrm[i];
    ushort4 g;
    float z = (gp.w < 0.0f) ? 0.0f : dot(sub(xyz(gp), cpos), cdir);
    g.x = f2h(gn.x); g.y = f2h(gn.y); g.z = f2h(gn.z); g.w = f2h(z);
    gb16[i] = g;
}

}  // namespace

extern "C" {

int rrh_atrous(rrh_ctx* c, const rrh_atrous_params* p, const rrh_image* input, const rrh_image* ga, const rrh_image* gbb,
               rrh_image* t0, rrh_image* t1, int* out_slot) {
    RRH_ARG(c && p && input && ga && t0 && t1 && out_slot, "rrh_atrous: null argument");
    RRH_ARG(p->mode == RRH_ATROUS_REFERENCE || p->mode == RRH_ATROUS_CONFIG, "rrh_atrous: bad mode");
    uint32_t W = input->w, H = input->h;
    RRH_ARG(ga->w == W && ga->h == H && t0->w == W && t0->h == H && t1->w == W && t1->h == H, "rrh_atrous: extent mismatch");
    uint32_t it = p->iterations < 1 ? 1 : (p->iterations > 7 ? 7 : p->iterations);
    dim3 grid((W + TX - 1) / TX, (H + TY - 1) / TY), block(TX * TY);
    rrh_image* slots[2] = {t0, t1};
    if (p->mode == RRH_ATROUS_REFERENCE) {
        RRH_ARG(gbb && gbb->w == W && gbb->h == H, "rrh_atrous: reference mode needs position and normal images");
        RRH_ARG(input->fmt == RRH_FMT_RGBA32F && ga->fmt == RRH_FMT_RGBA32F && gbb->fmt == RRH_FMT_RGBA32F &&
                    t0->fmt == RRH_FMT_RGBA32F && t1->fmt == RRH_FMT_RGBA32F, "rrh_atrous: reference mode uses RGBA32F");
        float sps = fmaxf(p->sigma_position * p->sigma_position, 1.0e-6f);
        float sns = fmaxf(p->sigma_normal * p->sigma_normal, 1.0e-6f);
        float sls = fmaxf(p->sigma_luminance * p->sigma_luminance, 1.0e-6f);
        for (uint32_t k = 0; k < it; ++k) {
            const float4* src = (const float4*)(k == 0 ? input->p : slots[(k - 1) & 1]->p);
            float4* dst = (float4*)slots[k & 1]->p;
            int st = 1 << k;
            if (st == 1)
                hipLaunchKernelGGL(k_atrous_ref_lds<1>, grid, block, 0, c->stream, (int)W, (int)H, sps, sns, sls, src,
                                   (const float4*)ga->p, (const float4*)gbb->p, dst);
            else if (st == 2)
                hipLaunchKernelGGL(k_atrous_ref_lds<2>, grid, block, 0, c->stream, (int)W, (int)H, sps, sns, sls, src,
                                   (const float4*)ga->p, (const float4*)gbb->p, dst);
            else
                hipLaunchKernelGGL(k_atrous_ref, grid, block, 0, c->stream, (int)W, (int)H, st, sps, sns, sls, src,
                                   (const float4*)ga->p, (const float4*)gbb->p, dst);
            RRH_LAUNCH_CHECK();
        }
    } else {
        RRH_ARG(input->fmt == RRH_FMT_RGBA16F && ga->fmt == RRH_FMT_RGBA16F && t0->fmt == RRH_FMT_RGBA16F &&
                    t1->fmt == RRH_FMT_RGBA16F, "rrh_atrous: config mode uses RGBA16F");
        for (uint32_t k = 0; k < it; ++k) {
            const ushort4* src = (const ushort4*)(k == 0 ? input->p : slots[(k - 1) & 1]->p);
            ushort4* dst = (ushort4*)slots[k & 1]->p;
            hipLaunchKernelGGL(k_atrous_cfg, grid, block, 0, c->stream, (int)W, (int)H, 1 << k, p->sigma_position,
                               p->sigma_normal, p->sigma_luminance, src, (const ushort4*)ga->p, dst);
            RRH_LAUNCH_CHECK();
        }
    }
    *out_slot = (int)((it - 1) & 1);
    return RRH_OK;
}

int rrh_atrous_prepare(rrh_ctx* c, const rrh_scene* s, const rrh_image* rad, const rrh_image* gpos,
                       const rrh_image* gnrm, rrh_image* rad16, rrh_image* gb16) {
    RRH_ARG(c && s && rad && gpos && gnrm && rad16 && gb16, "rrh_atrous_prepare: null argument");
    RRH_ARG(s->cam_set, "rrh_atrous_prepare: camera not set");
    uint32_t W = rad->w, H = rad->h;
    RRH_ARG(gpos->w == W && gnrm->w == W && rad16->w == W && gb16->w == W && gpos->h == H && gnrm->h == H &&
                rad16->h == H && gb16->h == H, "rrh_atrous_prepare: extent mismatch");
    RRH_ARG(rad->fmt == RRH_FMT_RGBA32F && rad16->fmt == RRH_FMT_RGBA16F && gb16->fmt == RRH_FMT_RGBA16F,
            "rrh_atrous_prepare: formats");
    const rrh_camera& cm = s->cam_host;
    dim3 grid((W + TX - 1) / TX, (H + TY - 1) / TY), block(TX * TY);
    hipLaunchKernelGGL(k_atrous_prepare, grid, block, 0, c->stream, (int)W, (int)H,
                       mk(cm.position[0], cm.position[1], cm.position[2]), mk(cm.direction[0], cm.direction[1], cm.direction[2]),
                       (const float4*)rad->p, (const float4*)gpos->p, (const float4*)gnrm->p, (ushort4*)rad16->p,
                       (ushort4*)gb16->p);
    RRH_LAUNCH_CHECK();
    return RRH_OK;
}

}  // extern "C"

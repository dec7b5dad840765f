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
// Each output pixel reads 25 taps whose footprint is shared with its neighbours: up to
// stride 8 every workgroup stages one window of its stride's sub-lattice in LDS (1.56 image
// reads per output) and the taps read LDS; at larger strides a dense tile reads its taps
// through L2.  With the exact transcendentals the per-tap arithmetic, not the memory
// traffic, bounds the kernels; config mode therefore keeps one exp per tap (its weights'
// arithmetic, which has no reference counterpart, is defined for that, as in the oracle).
#include <hip/hip_fp16.h>

#include "rrh_internal.h"

using namespace rrh;

namespace {

constexpr int TX = 16, TY = 16;

__device__ __forceinline__ float kw(int o) { return (o == -2 || o == 2) ? 1.0f : ((o == -1 || o == 1) ? 4.0f : 6.0f); }
__device__ __forceinline__ float luma601(f3 c) { return c.x * 0.299f + c.y * 0.587f + c.z * 0.114f; }
__device__ __forceinline__ int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

struct RefPx { f3 pos; float w; f3 nrm; f3 rgb; };

// Per-tap weight arithmetic.  Exact: correctly rounded division and square root and the
// library's own exp (wexp_rr for the weights) / log, bit-equal to the CPU restatement.  Fast: the hardware
// approximations (v_exp_f32, v_log_f32, v_rcp_f32, v_rsq_f32, about 1 ulp each); the taps'
// weights then differ in the last bits and the filtered pixel by far less than 1e-3.
template <bool FAST> struct WMath;
template <> struct WMath<false> {
    static __device__ __forceinline__ float ex(float x) { return exp_rr(x); }
    static __device__ __forceinline__ float wex(float x) { return wexp_rr(x); }  // the weights' exp (both modes)
    static __device__ __forceinline__ float div(float a, float b) { return a / b; }
    static __device__ __forceinline__ float rcp(float b) { return 1.0f / b; }
    static __device__ __forceinline__ f3 unit(f3 a) { return normalize(a); }
    static __device__ __forceinline__ float powp(float x, float e) { return (x > 0.0f) ? exp_rr(e * log_rr(x)) : 0.0f; }
};
template <> struct WMath<true> {
    static __device__ __forceinline__ float ex(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504f); }
    static __device__ __forceinline__ float wex(float x) { return ex(x); }
    static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    static __device__ __forceinline__ float rcp(float b) { return __builtin_amdgcn_rcpf(b); }
    static __device__ __forceinline__ f3 unit(f3 a) { return scale(a, __builtin_amdgcn_rsqf(dot(a, a))); }
    static __device__ __forceinline__ float powp(float x, float e) {
        return (x > 0.0f) ? __builtin_amdgcn_exp2f(e * __builtin_amdgcn_logf(x)) : 0.0f;
    }
};

// ---- sub-lattice tiling ---------------------------------------------------------------
// A workgroup's outputs are a TX x TY tile of the stride-S sub-lattice with residue (rx, ry):
// pixel (rx + S (lx0 + tx), ry + S (ly0 + ty)).  Every tap of those outputs lies on the same
// sub-lattice, clamping at the image edge included (lattice index k holds the image pixel
// clamp(rx + S k)), so the (TX + 4) x (TY + 4) lattice window is staged in LDS once (1.56
// loads per output) and all 25 taps read LDS.  The residue varies fastest
// over blockIdx.x, so the S x S workgroups of one lattice tile cover one dense image region.
constexpr int LW = TX + 4, LH = TY + 4;
#ifndef RRH_ATROUS_LAT_MAX
#define RRH_ATROUS_LAT_MAX 8
#endif
constexpr int kLatMaxStride = RRH_ATROUS_LAT_MAX;  // lattice staging up to this stride, dense tiles above

// Workgroups in XCD runs (xcd_block): the residues of one lattice tile and neighbouring
// tiles share cache lines, so they are kept on one XCD's L2.
__device__ __forceinline__ void xcd_tile(int& bx, int& by) {
    const uint32_t l = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y, true);
    bx = (int)(l % gridDim.x);
    by = (int)(l / gridDim.x);
}

struct LatTile {
    int rx, ry, lx0, ly0;
};
__device__ __forceinline__ LatTile lat_tile(int S, int bx, int by) {
    LatTile t;
    t.rx = bx % S;
    t.ry = by % S;
    t.lx0 = (bx / S) * TX;
    t.ly0 = (by / S) * TY;
    return t;
}
// image offset of window slot k (clamped to the image, as the taps are)
__device__ __forceinline__ size_t lat_src(const LatTile& t, int S, int W, int H, int k) {
    const int lx = k % LW, ly = k / LW;
    const int gx = clampi(t.rx + S * (t.lx0 - 2 + lx), 0, W - 1);
    const int gy = clampi(t.ry + S * (t.ly0 - 2 + ly), 0, H - 1);
    return (size_t)gy * W + gx;
}

// ---- reference mode -----------------------------------------------------------------
// LAT: the taps come from the staged sub-lattice window; otherwise every tap is read from
// global memory (through L2) for a dense TX x TY tile of pixels, which is faster at stride 16,
// where the lattice window's image reads (one pixel per 16) no longer share cache lines.
// The staged window holds each element's unit normal and luminance, computed once per element
// (about 1.6 per output) instead of once per tap: the same operations on the same inputs, so the
// same bits.  The weights' exponential is wexp_rr (the oracle's orc_wexpf).
template <bool FAST, bool LAT>
__global__ void __launch_bounds__(TX* TY) k_atrous_ref(int W, int H, int S, float sps, float sns, float sls,
                                                      const float4* __restrict__ in, const float4* __restrict__ gpos,
                                                      const float4* __restrict__ gnrm, float4* __restrict__ out) {
    __shared__ float4 s_pos[LAT ? LW * LH : 1];
    __shared__ float4 s_nrm[LAT ? LW * LH : 1];  // unit normal, luminance of the radiance in w
    __shared__ float4 s_rad[LAT ? LW * LH : 1];
    int bx, by;
    xcd_tile(bx, by);
    const LatTile t = LAT ? lat_tile(S, bx, by) : LatTile{0, 0, 0, 0};
    if constexpr (LAT) {
        for (int k = threadIdx.x; k < LW * LH; k += TX * TY) {
            const size_t g = lat_src(t, S, W, H, k);
            const float4 n = gnrm[g], r = in[g];  // (the normal's w is unused)
            const f3 u = WMath<FAST>::unit(xyz(n));
            s_pos[k] = gpos[g];
            s_nrm[k] = make_float4(u.x, u.y, u.z, luma601(xyz(r)));
            s_rad[k] = r;
        }
        __syncthreads();
    }
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    const int x = LAT ? t.rx + S * (t.lx0 + tx) : bx * TX + tx;
    const int y = LAT ? t.ry + S * (t.ly0 + ty) : by * TY + ty;
    if (x >= W || y >= H) return;
    const int lc = (ty + 2) * LW + (tx + 2);
    size_t i = (size_t)y * W + x;
    // tap (dx, dy) in lattice steps: its position, unit normal, radiance and luminance; off the
    // lattice a position with w < 0 (no surface) leaves the normal and luminance unset
    auto tap = [&](int dx, int dy, float4& p, f3& n, float4& r, float& lum) {
        if constexpr (LAT) {  // LDS reads, unconditional (no branch per tap)
            const int l = lc + dy * LW + dx;
            const float4 nl = s_nrm[l];
            p = s_pos[l]; n = xyz(nl); r = s_rad[l]; lum = nl.w;
        } else {
            const size_t j = (size_t)clampi(y + dy * S, 0, H - 1) * W + clampi(x + dx * S, 0, W - 1);
            p = gpos[j];
            r = in[j];
            if (p.w < 0.0f) return;
            n = WMath<FAST>::unit(xyz(gnrm[j])); lum = luma601(xyz(r));
        }
    };
    float4 cp, cr;
    f3 cn;
    float cl;
    tap(0, 0, cp, cn, cr, cl);
    if (cp.w < 0.0f) { out[i] = cr; return; }
    f3 acc = mk(0, 0, 0);
    float wsum = 0.0f;
    for (int ty2 = -2; ty2 <= 2; ++ty2) {
        for (int tx2 = -2; tx2 <= 2; ++tx2) {
            float4 sp, sr4;
            f3 sn;
            float sl;
            tap(tx2, ty2, sp, sn, sr4, sl);
            if (sp.w < 0.0f) continue;
            f3 sr = xyz(sr4);
            float bw = (kw(tx2) * kw(ty2)) / 256.0f;
            f3 pd = sub(xyz(sp), xyz(cp));
            float pw = WMath<FAST>::wex(-WMath<FAST>::div(dot(pd, pd), sps));
            float nd = fmaxf(0.0f, 1.0f - dot(cn, sn));
            float nw = WMath<FAST>::wex(-WMath<FAST>::div(nd, sns));
            float ld = sl - cl;
            float lw = WMath<FAST>::wex(-WMath<FAST>::div(ld * ld, sls));
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


// The taps come from the staged lattice window; the centre's 3x3 variance and its depth
// gradient (stride 1, off the lattice) are read from global memory.
// Per tap: one exp (wexp_rr, a fused-multiply-add polynomial); the normal dot product, the
// exp's argument -|dl| / phi_l - wz and the weighted sums are fused multiply-adds.  The normal
// weight max(0, n.n)^sn is nsq squarings when sn = 2^nsq (nsq < 0: exp(sn log x)); the luminance
// and depth terms multiply by reciprocals formed once per pixel (the depth term's 13: tap -o
// shares tap o's denominator exactly).
template <bool FAST, bool LAT>
__global__ void __launch_bounds__(TX* TY) k_atrous_cfg(int W, int H, int S, float sz, float sn, int nsq, float sl,
                                                      const ushort4* __restrict__ in, const ushort4* __restrict__ gb,
                                                      ushort4* __restrict__ out) {
    // the lattice window staged once in fp32 (the fp16 inputs convert exactly) with each
    // element's luminance, so the taps neither convert nor recompute it
    __shared__ float4 s_gb[LAT ? LW * LH : 1];   // normal, linear depth
    __shared__ float4 s_in[LAT ? LW * LH : 1];   // rgb, variance
    __shared__ float s_lum[LAT ? LW * LH : 1];   // lum709(rgb)
    int bx, by;
    xcd_tile(bx, by);
    const LatTile t = LAT ? lat_tile(S, bx, by) : LatTile{0, 0, 0, 0};
    if constexpr (LAT) {
        for (int k = threadIdx.x; k < LW * LH; k += TX * TY) {
            const size_t g = lat_src(t, S, W, H, k);
            const ushort4 a = gb[g], b = in[g];
            s_gb[k] = make_float4(h2f(a.x), h2f(a.y), h2f(a.z), h2f(a.w));
            const f3 c = mk(h2f(b.x), h2f(b.y), h2f(b.z));
            s_in[k] = make_float4(c.x, c.y, c.z, h2f(b.w));
            s_lum[k] = lum709(c);
        }
        __syncthreads();
    }
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    const int x = LAT ? t.rx + S * (t.lx0 + tx) : bx * TX + tx;
    const int y = LAT ? t.ry + S * (t.ly0 + ty) : by * TY + ty;
    if (x >= W || y >= H) return;
    const int lc = (ty + 2) * LW + (tx + 2);
    size_t i = (size_t)y * W + x;
    // tap (dx, dy): normal + depth, rgb + variance, luminance
    auto tap = [&](int dx, int dy, float4& g, float4& r, float& l) {
        if constexpr (LAT) {
            const int k = lc + dy * LW + dx;
            g = s_gb[k]; r = s_in[k]; l = s_lum[k];
        } else {
            const size_t j = (size_t)clampi(y + dy * S, 0, H - 1) * W + clampi(x + dx * S, 0, W - 1);
            const ushort4 a = gb[j], b = in[j];
            g = make_float4(h2f(a.x), h2f(a.y), h2f(a.z), h2f(a.w));
            r = make_float4(h2f(b.x), h2f(b.y), h2f(b.z), h2f(b.w));
            l = lum709(xyz(r));
        }
    };
    float4 cg, ci;
    float lp;
    tap(0, 0, cg, ci, lp);
    float zp = cg.w;
    if (!(zp > 0.0f)) { out[i] = in[i]; return; }
    f3 np = xyz(cg);
    const float gk[3] = {0.25f, 0.5f, 0.25f};
    float var = 0.0f;
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            int sx = clampi(x + dx, 0, W - 1), sy = clampi(y + dy, 0, H - 1);
            var += gk[dx + 1] * gk[dy + 1] * h2f(in[(size_t)sy * W + sx].w);
        }
    float phil = sl * sqrtf(fmaxf(0.0f, var)) + 1e-10f;
    const float inv_phil = WMath<FAST>::rcp(phil);
    int xm = x > 0 ? x - 1 : 0, xp = x < W - 1 ? x + 1 : W - 1;
    int ym = y > 0 ? y - 1 : 0, yp = y < H - 1 ? y + 1 : H - 1;
    float dzdx = (h2f(gb[(size_t)y * W + xp].w) - h2f(gb[(size_t)y * W + xm].w)) * 0.5f;
    float dzdy = (h2f(gb[(size_t)yp * W + x].w) - h2f(gb[(size_t)ym * W + x].w)) * 0.5f;
    const float gx = sz * dzdx * (float)S, gy = sz * dzdy * (float)S;
    float rden[13];  // taps k = 5 (ty + 2) + (tx + 2) <= 12; tap 24 - k mirrors tap k
#pragma unroll
    for (int k = 0; k < 13; ++k)
        rden[k] = WMath<FAST>::rcp(fabsf(gx * (float)(k % 5 - 2) + gy * (float)(k / 5 - 2)) + 1e-3f);
    f3 acc = mk(0, 0, 0);
    float accv = 0.0f, wsum = 0.0f;
#pragma unroll
    for (int ty2 = -2; ty2 <= 2; ++ty2) {
#pragma unroll
        for (int tx2 = -2; tx2 <= 2; ++tx2) {
            float4 sg, si;
            float lq;
            tap(tx2, ty2, sg, si, lq);
            float zq = sg.w;
            if (!(zq > 0.0f)) continue;
            f3 nq = xyz(sg);
            f3 cq = xyz(si);
            float vq = si.w;
            float bw = (kw(tx2) * kw(ty2)) / 256.0f;
            const int k = 5 * (ty2 + 2) + (tx2 + 2);
            float wz = fabsf(zp - zq) * rden[k <= 12 ? k : 24 - k];
            float wn = fmaxf(0.0f, __builtin_fmaf(np.z, nq.z, __builtin_fmaf(np.y, nq.y, np.x * nq.x)));
            if (nsq >= 0) {
                for (int q = 0; q < nsq; ++q) wn = wn * wn;
            } else {
                wn = WMath<FAST>::powp(wn, sn);
            }
            float w = bw * WMath<FAST>::wex(__builtin_fmaf(-fabsf(lp - lq), inv_phil, -wz)) * wn;
            acc.x = __builtin_fmaf(cq.x, w, acc.x);
            acc.y = __builtin_fmaf(cq.y, w, acc.y);
            acc.z = __builtin_fmaf(cq.z, w, acc.z);
            accv = __builtin_fmaf(vq, w * w, accv);
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

// S x S workgroups per TX x TY tile of each residue's sub-lattice
dim3 lat_grid(uint32_t W, uint32_t H, int S) {
    const uint32_t lw = (W + S - 1) / S, lh = (H + S - 1) / S;
    return dim3(S * ((lw + TX - 1) / TX), S * ((lh + TY - 1) / TY));
}

}  // namespace

extern "C" {

int rrh_atrous(rrh_ctx* c, const rrh_atrous_params* p, const rrh_image* input, const rrh_image* ga, const rrh_image* gbb,
               rrh_image* t0, rrh_image* t1, int* out_slot) {
    RRH_ARG(c && p && input && ga && t0 && t1 && out_slot, "rrh_atrous: null argument");
    RRH_ARG(p->mode == RRH_ATROUS_REFERENCE || p->mode == RRH_ATROUS_CONFIG, "rrh_atrous: bad mode");
    RRH_ARG((p->flags & ~RRH_ATROUS_FAST_MATH) == 0u, "rrh_atrous: unknown flags");
    const bool fast = (p->flags & RRH_ATROUS_FAST_MATH) != 0u;
    uint32_t W = input->w, H = input->h;
    RRH_ARG(ga->w == W && ga->h == H && t0->w == W && t0->h == H && t1->w == W && t1->h == H, "rrh_atrous: extent mismatch");
    uint32_t it = p->iterations < 1 ? 1 : (p->iterations > 7 ? 7 : p->iterations);
    const dim3 block(TX * TY), dense((W + TX - 1) / TX, (H + TY - 1) / TY);
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
            const int S = 1 << k;
            const bool lat = S <= kLatMaxStride;
            auto kern = fast ? (lat ? k_atrous_ref<true, true> : k_atrous_ref<true, false>)
                             : (lat ? k_atrous_ref<false, true> : k_atrous_ref<false, false>);
            hipLaunchKernelGGL(kern, lat ? lat_grid(W, H, S) : dense, block, 0, c->stream, (int)W, (int)H, S, sps, sns, sls,
                               src, (const float4*)ga->p, (const float4*)gbb->p, dst);
            RRH_LAUNCH_CHECK();
        }
    } else {
        RRH_ARG(input->fmt == RRH_FMT_RGBA16F && ga->fmt == RRH_FMT_RGBA16F && t0->fmt == RRH_FMT_RGBA16F &&
                    t1->fmt == RRH_FMT_RGBA16F, "rrh_atrous: config mode uses RGBA16F");
        int nsq = -1;  // sigma_normal = 2^nsq: the normal weight by squaring
        for (int k = 0; k <= 10; ++k)
            if (p->sigma_normal == (float)(1 << k)) { nsq = k; break; }
        for (uint32_t k = 0; k < it; ++k) {
            const ushort4* src = (const ushort4*)(k == 0 ? input->p : slots[(k - 1) & 1]->p);
            ushort4* dst = (ushort4*)slots[k & 1]->p;
            const int S = 1 << k;
            const bool lat = S <= kLatMaxStride;
            auto kern = fast ? (lat ? k_atrous_cfg<true, true> : k_atrous_cfg<true, false>)
                             : (lat ? k_atrous_cfg<false, true> : k_atrous_cfg<false, false>);
            hipLaunchKernelGGL(kern, lat ? lat_grid(W, H, S) : dense, block, 0, c->stream, (int)W, (int)H, S, p->sigma_position,
                               p->sigma_normal, nsq, p->sigma_luminance, src, (const ushort4*)ga->p, dst);
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

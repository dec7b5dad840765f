// rrh_mlp.hip -- learned path guide: a small MLP that maps (position, normal, outgoing
// direction) to an 8-lobe vMF mixture, trained online on the path tracer's own vertices.
// The reference has no learned guide (its guide is the single-lobe grid of
// GuidingGrid.slang); the network and loss are defined in include/rrh.h.
//
// Training step = k_mlp_wsum (sum of sample weights) -> k_mlp_train (persistent blocks:
// per 64-sample chunk, forward through five BF16 MFMA layers, vMF-mixture NLL head,
// backward deltas, then the block's weight-gradient tiles, kept in registers across
// chunks) -> k_mlp_reduce (deterministic sum of the per-block partials, non-finite
// check) -> k_mlp_adam (FP32 master weights) -> k_mlp_pack (BF16 copies for the next
// step) -> k_mlp_finish (step count, dynamic loss scale).
//
// MFMA: v_mfma_f32_16x16x32_bf16.  A (16x32): lane l holds row l%16, k = 8*(l/16)..+7;
// B (32x16): column l%16, same k; D (16x16 f32): column l%16, rows 4*(l/16)..+3.
#include "rrh_internal.h"
#include "rrh_guide_dev.h"

using namespace rrh;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int NI = 78, KI = 96, NH = 64, NO = 40, NOP = 48, NK = 8;
constexpr int NPARAM = RRH_MLP_PARAMS;
constexpr int kBlockSamples = 64;   // 4 waves x 16 samples
constexpr int kTiles = 80;          // weight-gradient tiles (16x16) over all layers

__host__ __device__ constexpr int in_of(int l) { return l == 0 ? NI : NH; }
__host__ __device__ constexpr int out_of(int l) { return l == 4 ? NO : NH; }
__host__ __device__ constexpr int off_w(int l) {
    int o = 0;
    for (int i = 0; i < l; ++i) o += in_of(i) * out_of(i) + out_of(i);
    return o;
}
__host__ __device__ constexpr int off_b(int l) { return off_w(l) + in_of(l) * out_of(l); }
static_assert(off_w(5) == NPARAM, "parameter count");

// bf16 device copies: W_l [Opad][Kpad] (forward B operand), WT_l [64][64] (backward B operand)
constexpr int kWOff[5] = {0, 64 * 96, 64 * 96 + 4096, 64 * 96 + 8192, 64 * 96 + 12288};
constexpr int kWTOff = 64 * 96 + 12288 + 48 * 64;  // WT1..WT4, 4096 each
constexpr int kPackElems = kWTOff + 4 * 4096;

struct MlpDev {
    const __bf16* pk;  // packed bf16 weights
    const float* P;    // fp32 master parameters (biases are read from here)
    float bmin[3], bext[3];
};

__device__ __forceinline__ const __bf16* Wl(const MlpDev& M, int l) { return M.pk + kWOff[l]; }
__device__ __forceinline__ const __bf16* WTl(const MlpDev& M, int l) { return M.pk + kWTOff + (l - 1) * 4096; }

// ---- inputs --------------------------------------------------------------------------
struct Sample { f3 p, n, wo, dir; float w; };

struct HostSrc {  // arrays uploaded from the host (3 floats per vector)
    const float *pos, *nrm, *wo, *dir, *w;
    uint32_t n;
    __device__ uint32_t count() const { return n; }
    __device__ bool get(uint32_t i, Sample& s) const {
        if (i >= n) return false;
        s.p = mk(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        s.n = mk(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
        s.wo = mk(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]);
        s.dir = dir ? mk(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]) : mk(0, 0, 1);
        s.w = w ? w[i] : 0.0f;
        return true;
    }
};
struct TrainSrc {  // samples finalised from path-vertex records: t0 (pos, w) t1 (dir, pdf) t2 (n) t3 (wo)
    const float4* tr;
    const uint32_t* cnt;
    uint32_t cap, batch;
    __device__ uint32_t count() const { return min(min(*cnt, cap), batch); }
    __device__ bool get(uint32_t i, Sample& s) const {
        if (i >= count()) return false;
        float4 t0 = tr[(size_t)i * 4], t1 = tr[(size_t)i * 4 + 1], t2 = tr[(size_t)i * 4 + 2], t3 = tr[(size_t)i * 4 + 3];
        s.p = xyz(t0); s.w = t0.w; s.dir = xyz(t1); s.n = xyz(t2); s.wo = xyz(t3);
        return true;
    }
};

// ---- encoding: features 0..71 = sin/cos(2^j pi u_c), 72..74 n, 75..77 wo, 78..95 zero ----
// 4 lanes per sample; lane quarter q writes features [24q, 24q + 24)
__device__ __forceinline__ void encode_row(const MlpDev& M, const Sample& s, bool valid, int q, __bf16* row) {
    if (q < 3) {
        float pc = (q == 0) ? s.p.x : (q == 1 ? s.p.y : s.p.z);
        float u = valid ? (pc - M.bmin[q]) / M.bext[q] : 0.0f;
        u = clampf(u, 0.0f, 1.0f);
        for (int j = 0; j < 12; ++j) {
            float t = u * (float)(1 << j);               // exact scaling by 2^j
            float tm = t - 2.0f * floorf(t * 0.5f);      // t mod 2, exact for t < 2^24
            float sn, cs;
            sincos_rr(3.14159265358979323846f * tm, &sn, &cs);
            row[q * 24 + 2 * j] = (__bf16)(valid ? sn : 0.0f);
            row[q * 24 + 2 * j + 1] = (__bf16)(valid ? cs : 0.0f);
        }
    } else {
        float v[6] = {s.n.x, s.n.y, s.n.z, s.wo.x, s.wo.y, s.wo.z};
        for (int k = 0; k < 24; ++k) row[72 + k] = (__bf16)((valid && k < 6) ? v[k] : 0.0f);
    }
}

// ---- one 16-sample wave through a dense layer: out = A W^T + b ----------------------
template <int K, int NT>
__device__ __forceinline__ void layer_fwd(const __bf16* A, const __bf16* W, const float* bias, int nout, bool relu,
                                          __bf16* H, float* O) {
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kk = 0; kk < K / 32; ++kk) {
            bf16x8 a = *reinterpret_cast<const bf16x8*>(A + r * K + kk * 32 + g * 8);
            bf16x8 b = *reinterpret_cast<const bf16x8*>(W + (nt * 16 + r) * K + kk * 32 + g * 8);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
        }
        const int o = nt * 16 + r;
        const float bo = (o < nout) ? bias[o] : 0.0f;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int s = 4 * g + j;
            float z = acc[j] + bo;
            if (relu) H[s * 64 + o] = (__bf16)fmaxf(z, 0.0f);
            else O[s * NOP + o] = (o < nout) ? z : 0.0f;
        }
    }
}

// delta through a layer: Dout = (Din W) * (Hmask > 0); Din, Hmask, Dout are [16][64]
__device__ __forceinline__ void layer_bwd(const __bf16* Din, const __bf16* WT, const __bf16* Hm, __bf16* Dout) {
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
            bf16x8 a = *reinterpret_cast<const bf16x8*>(Din + r * 64 + kk * 32 + g * 8);
            bf16x8 b = *reinterpret_cast<const bf16x8*>(WT + (nt * 16 + r) * 64 + kk * 32 + g * 8);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
        }
        const int i = nt * 16 + r;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int s = 4 * g + j;
            float m = ((float)Hm[s * 64 + i] > 0.0f) ? acc[j] : 0.0f;
            Dout[s * 64 + i] = (__bf16)m;
        }
    }
}

// ---- head: lobes from the 40 raw outputs, and the NLL gradient -----------------------
__device__ __forceinline__ float softplus_f(float d) { return d > 20.0f ? d : log_rr(1.0f + exp_rr(d)); }
__device__ __forceinline__ float sigmoid_f(float d) { return 1.0f / (1.0f + exp_rr(-d)); }

__device__ __forceinline__ void head_lobes(const float* o, float* lobes) {
    float emax = -3.0e38f;
    for (int k = 0; k < NK; ++k) emax = fmaxf(emax, o[k * 5 + 4]);
    float z = 0.0f;
    for (int k = 0; k < NK; ++k) z += exp_rr(o[k * 5 + 4] - emax);
    for (int k = 0; k < NK; ++k) {
        const float* v = o + k * 5;
        float len = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + 1e-12f);
        lobes[k * 5 + 0] = v[0] / len;
        lobes[k * 5 + 1] = v[1] / len;
        lobes[k * 5 + 2] = v[2] / len;
        lobes[k * 5 + 3] = fminf(0.1f + softplus_f(v[3]), 1000.0f);
        lobes[k * 5 + 4] = exp_rr(v[4] - emax) / z;
    }
}
__device__ __forceinline__ float vmf_val(float kap, float c) {
    return kap * exp_rr(kap * (c - 1.0f)) / (6.28318530717958647692f * (1.0f - exp_rr(-2.0f * kap)));
}
// returns the sample's loss term; writes dL/dout (times `gscale`) into g[40]
__device__ __forceinline__ float head_grad(const float* o, f3 d, float wn, float gscale, float* g) {
    float lob[NK * 5];
    head_lobes(o, lob);
    float f[NK], c[NK], q = 0.0f;
    for (int k = 0; k < NK; ++k) {
        c[k] = lob[k * 5] * d.x + lob[k * 5 + 1] * d.y + lob[k * 5 + 2] * d.z;
        f[k] = vmf_val(lob[k * 5 + 3], c[k]);
        q += lob[k * 5 + 4] * f[k];
    }
    float loss = -wn * log_rr(q + 1e-20f);
    float dq = -wn / (q + 1e-20f) * gscale;
    for (int k = 0; k < NK; ++k) {
        const float* v = o + k * 5;
        float pik = lob[k * 5 + 4], kap = lob[k * 5 + 3];
        g[k * 5 + 4] = dq * pik * (f[k] - q);
        float e2 = exp_rr(-2.0f * kap);
        float dlogf = 1.0f / kap - 2.0f * e2 / (1.0f - e2) + (c[k] - 1.0f);
        float raw = 0.1f + softplus_f(v[3]);
        g[k * 5 + 3] = (raw > 1000.0f) ? 0.0f : dq * pik * f[k] * dlogf * sigmoid_f(v[3]);
        float len = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + 1e-12f);
        float mu0 = lob[k * 5], mu1 = lob[k * 5 + 1], mu2 = lob[k * 5 + 2];
        float s = dq * pik * f[k] * kap;
        float g0 = s * d.x, g1 = s * d.y, g2 = s * d.z;
        float mg = mu0 * g0 + mu1 * g1 + mu2 * g2;
        g[k * 5 + 0] = (g0 - mu0 * mg) / len;
        g[k * 5 + 1] = (g1 - mu1 * mg) / len;
        g[k * 5 + 2] = (g2 - mu2 * mg) / len;
    }
    return loss;
}

// ---- shared memory of a 4-wave block (64 samples) -------------------------------------
// Training keeps every activation and delta of the chunk for the weight gradients; the raw
// outputs alias D[0..1], which are written only after the head has consumed them.
struct __align__(16) TrainSmem {
    __bf16 X[kBlockSamples * KI];        // encoded inputs
    __bf16 H[4][kBlockSamples * 64];     // post-ReLU activations of layers 0..3
    __bf16 D[5][kBlockSamples * 64];     // dL/dz of layers 0..4 (D[4]: output layer, 40 used)
    float red[8];
    __device__ float* O(int w) { return reinterpret_cast<float*>(&D[0][0]) + w * 16 * NOP; }
};
static_assert(4 * 16 * NOP * 4 <= 2 * kBlockSamples * 64 * 2, "raw outputs fit in D[0..1]");
struct __align__(16) InferSmem {
    __bf16 X[kBlockSamples * KI];
    __bf16 H[2][kBlockSamples * 64];     // ping-pong
    float O[4 * 16 * NOP];
};

template <class Src>
__device__ void forward_chunk(const MlpDev& M, const Src& src, uint32_t base, __bf16* X, __bf16* const* H, float* Ow,
                              Sample& mine, bool& valid) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // encode: 4 lanes per sample
    {
        const int sl = lane >> 2, q = lane & 3;
        Sample s;
        bool v = src.get(base + w * 16 + sl, s);
        if (!v) { s.p = s.n = s.wo = s.dir = mk(0, 0, 0); s.w = 0.0f; }
        encode_row(M, s, v, q, X + (w * 16 + sl) * KI);
        if (q == 0) { mine = s; valid = v; }
    }
    __syncthreads();
    layer_fwd<KI, 4>(X + w * 16 * KI, Wl(M, 0), M.P + off_b(0), NH, true, H[0] + w * 16 * 64, nullptr);
    __syncthreads();
    for (int l = 1; l < 4; ++l) {
        layer_fwd<64, 4>(H[l - 1] + w * 16 * 64, Wl(M, l), M.P + off_b(l), NH, true, H[l] + w * 16 * 64, nullptr);
        __syncthreads();
    }
    layer_fwd<64, 3>(H[3] + w * 16 * 64, Wl(M, 4), M.P + off_b(4), NO, false, nullptr, Ow);
    __syncthreads();
}

__device__ __forceinline__ void tile_of(int t, int& l, int& ot, int& it) {
    if (t < 20) { l = 0; ot = t / 5; it = t % 5; }
    else if (t < 68) { int u = t - 20; l = 1 + u / 16; ot = (u % 16) / 4; it = u % 4; }
    else { int u = t - 68; l = 4; ot = u / 4; it = u % 4; }
}

template <class Src>
__global__ void __launch_bounds__(256) k_mlp_train(MlpDev M, Src src, uint32_t nsamples, const float* __restrict__ wsum,
                                                   const float* __restrict__ scale_p, float* __restrict__ partial,
                                                   float* __restrict__ partial_loss) {
    __shared__ TrainSmem sm;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    __bf16* const Hs[4] = {sm.H[0], sm.H[1], sm.H[2], sm.H[3]};
    const float inv_w = 1.0f / fmaxf(*wsum, 1e-30f);
    const float gscale = *scale_p;
    f32x4 acc[kTiles / 4];
#pragma unroll
    for (int t = 0; t < kTiles / 4; ++t) acc[t] = {0.0f, 0.0f, 0.0f, 0.0f};
    float bacc0 = 0.0f, bacc1 = 0.0f, loss_acc = 0.0f;
    const uint32_t n = min(src.count(), nsamples);
    for (uint32_t base = blockIdx.x * kBlockSamples; base < n; base += gridDim.x * kBlockSamples) {
        Sample mine;
        bool valid = false;
        forward_chunk(M, src, base, sm.X, Hs, sm.O(w), mine, valid);
        // head: 16 lanes per wave, one sample each (lane 4*sl holds sample sl after encode)
        {
            const int sl = lane >> 2;
            float gout[NO];
            bool act = (lane & 3) == 0;
            float lw = 0.0f;
            if (act) {
                const float* o = sm.O(w) + sl * NOP;
                float wn = valid ? mine.w * inv_w : 0.0f;
                if (wn > 0.0f) lw = head_grad(o, mine.dir, wn, gscale, gout);
                else for (int k = 0; k < NO; ++k) gout[k] = 0.0f;
                __bf16* Drow = sm.D[4] + (w * 16 + sl) * 64;
                for (int k = 0; k < 64; ++k) Drow[k] = (__bf16)(k < NO ? gout[k] : 0.0f);
            }
            loss_acc += lw;
        }
        __syncthreads();
        // deltas: D3 = (D4 W4) * (H3 > 0) ... D0 = (D1 W1) * (H0 > 0)
        for (int l = 4; l >= 1; --l) {
            layer_bwd(sm.D[l] + w * 16 * 64, WTl(M, l), sm.H[l - 1] + w * 16 * 64, sm.D[l - 1] + w * 16 * 64);
            __syncthreads();
        }
        // weight-gradient tiles: dW_l[o][i] += sum_s D_l[s][o] A_l[s][i]  (K = the 64 samples)
#pragma unroll
        for (int tt = 0; tt < kTiles / 4; ++tt) {
            int l, ot, it;
            tile_of(tt * 4 + w, l, ot, it);
            const __bf16* Dl = sm.D[l];
            const __bf16* Al = (l == 0) ? sm.X : sm.H[l - 1];
            const int alen = (l == 0) ? KI : 64;
#pragma unroll
            for (int kk = 0; kk < 2; ++kk) {
                bf16x8 a, b;
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int s = kk * 32 + g * 8 + j;
                    a[j] = Dl[s * 64 + ot * 16 + r];
                    b[j] = Al[s * alen + it * 16 + r];
                }
                acc[tt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[tt], 0, 0, 0);
            }
        }
        // bias gradients: thread t owns bias t of layers 0..3, threads < 40 also layer 4
        {
            const int t = threadIdx.x, l = t >> 6, o = t & 63;
            float s0 = 0.0f, s1 = 0.0f;
            for (int s = 0; s < kBlockSamples; ++s) {
                s0 += (float)sm.D[l][s * 64 + o];
                if (t < NO) s1 += (float)sm.D[4][s * 64 + t];
            }
            bacc0 += s0;
            bacc1 += s1;
        }
        __syncthreads();
    }
    // write this block's partial gradient (each parameter written once per block)
    float* part = partial + (size_t)blockIdx.x * NPARAM;
#pragma unroll
    for (int tt = 0; tt < kTiles / 4; ++tt) {
        int l, ot, it;
        tile_of(tt * 4 + w, l, ot, it);
        const int i = it * 16 + r;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int o = ot * 16 + 4 * g + j;
            if (o < out_of(l) && i < in_of(l)) part[off_w(l) + o * in_of(l) + i] = acc[tt][j];
        }
    }
    {
        const int t = threadIdx.x, l = t >> 6, o = t & 63;
        part[off_b(l) + o] = bacc0;
        if (t < NO) part[off_b(4) + t] = bacc1;
    }
    // block loss
    float v = loss_acc;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (lane == 0) sm.red[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) partial_loss[blockIdx.x] = sm.red[0] + sm.red[1] + sm.red[2] + sm.red[3];
}

template <class Src>
__global__ void __launch_bounds__(256) k_mlp_infer(MlpDev M, Src src, uint32_t nsamples, float* __restrict__ lobes) {
    __shared__ InferSmem sm;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __bf16* const Hs[4] = {sm.H[0], sm.H[1], sm.H[0], sm.H[1]};
    const uint32_t n = min(src.count(), nsamples);
    for (uint32_t base = blockIdx.x * kBlockSamples; base < n; base += gridDim.x * kBlockSamples) {
        Sample mine;
        bool valid = false;
        forward_chunk(M, src, base, sm.X, Hs, sm.O + w * 16 * NOP, mine, valid);
        const int sl = lane >> 2;
        if ((lane & 3) == 0 && valid) {
            float lob[NK * 5];
            head_lobes(sm.O + w * 16 * NOP + sl * NOP, lob);
            float* dst = lobes + (size_t)(base + w * 16 + sl) * (NK * 5);
            for (int k = 0; k < NK * 5; ++k) dst[k] = lob[k];
        }
        __syncthreads();
    }
}

__global__ void k_mlp_wsum_host(const float* __restrict__ w, uint32_t n, float* __restrict__ out) {
    float v = 0.0f;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v += w[i];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, v);
}
__global__ void k_mlp_wsum_train(const float4* __restrict__ tr, const uint32_t* __restrict__ cnt, uint32_t cap,
                                 uint32_t batch, float* __restrict__ out) {
    uint32_t n = min(min(*cnt, cap), batch);
    float v = 0.0f;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v += tr[(size_t)i * 4].w;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, v);
}

// grad = sum over blocks / loss scale (fixed order: deterministic); flags non-finite values
__global__ void k_mlp_reduce(const float* __restrict__ partial, const float* __restrict__ partial_loss, uint32_t nb,
                             const float* __restrict__ scale_p, float* __restrict__ grad, float* __restrict__ loss,
                             uint32_t* __restrict__ bad) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    float inv = 1.0f / *scale_p;
    if (i < (uint32_t)NPARAM) {
        float s = 0.0f;
        for (uint32_t b = 0; b < nb; ++b) s += partial[(size_t)b * NPARAM + i];
        s *= inv;
        if (!isfinite(s)) atomicOr(bad, 1u);
        grad[i] = s;
    }
    if (i == 0) {
        float l = 0.0f;
        for (uint32_t b = 0; b < nb; ++b) l += partial_loss[b];
        *loss = l;
    }
}

struct AdamArgs { float lr, b1, b2, eps; };

// Adam (Kingma & Ba 2015) on the FP32 master weights; skipped when the step overflowed
__global__ void k_mlp_adam(float* __restrict__ P, const float* __restrict__ grad, float* __restrict__ m,
                           float* __restrict__ v, const uint32_t* __restrict__ state, const uint32_t* __restrict__ bad,
                           AdamArgs A) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint32_t)NPARAM || *bad) return;
    float t = (float)(state[0] + 1u);
    float c1 = 1.0f - exp_rr(t * log_rr(A.b1)), c2 = 1.0f - exp_rr(t * log_rr(A.b2));
    float g = grad[i];
    float mi = A.b1 * m[i] + (1.0f - A.b1) * g;
    float vi = A.b2 * v[i] + (1.0f - A.b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    P[i] = P[i] - A.lr * (mi / c1) / (sqrtf(vi / c2) + A.eps);
}

// state[0] = Adam steps, state[1] = good steps since the last scale change; scale halves on
// overflow and doubles after 2000 good steps
__global__ void k_mlp_finish(uint32_t* __restrict__ state, uint32_t* __restrict__ bad, float* __restrict__ scale) {
    if (threadIdx.x != 0) return;
    if (*bad) {
        *scale = fmaxf(*scale * 0.5f, 1.0f);
        state[1] = 0;
    } else {
        state[0] += 1;
        state[1] += 1;
        if (state[1] >= 2000u) { *scale = fminf(*scale * 2.0f, 65536.0f); state[1] = 0; }
    }
    *bad = 0;
}

// bf16 copies: W_l [Opad][Kpad] and WT_l [64][64] (l >= 1), zero padded
__global__ void k_mlp_pack(const float* __restrict__ P, __bf16* __restrict__ pk) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= kPackElems) return;
    float val = 0.0f;
    if (e < kWTOff) {
        int l = 4;
        while (l > 0 && e < kWOff[l]) --l;
        int local = e - kWOff[l];
        int kp = (l == 0) ? KI : 64;
        int o = local / kp, i = local % kp;
        if (o < out_of(l) && i < in_of(l)) val = P[off_w(l) + o * in_of(l) + i];
    } else {
        int local = e - kWTOff;
        int l = 1 + local / 4096;
        int i = (local % 4096) / 64, o = local % 64;
        if (o < out_of(l) && i < in_of(l)) val = P[off_w(l) + o * in_of(l) + i];
    }
    pk[e] = (__bf16)val;
}

uint32_t pcg_host(uint32_t s) {
    uint32_t state = s * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

}  // namespace

struct rrh_mlp {
    rrh_ctx* ctx = nullptr;
    rrh_mlp_params p{};
    DevBuf P, m, v, grad, pk, partial, ploss, scal;  // scal: wsum, scale, loss (floats) + state, bad (u32)
    uint32_t nb = 0;
    MlpDev dev() const {
        MlpDev d;
        d.pk = (const __bf16*)pk.p;
        d.P = (const float*)P.p;
        for (int k = 0; k < 3; ++k) { d.bmin[k] = p.bounds_min[k]; d.bext[k] = p.bounds_max[k] - p.bounds_min[k]; }
        return d;
    }
    float* wsum() const { return (float*)scal.p; }
    float* scale() const { return (float*)scal.p + 1; }
    float* loss() const { return (float*)scal.p + 2; }
    uint32_t* state() const { return (uint32_t*)scal.p + 4; }
    uint32_t* bad() const { return (uint32_t*)scal.p + 6; }
};

namespace {

int mlp_pack(rrh_ctx* c, rrh_mlp* m) {
    hipLaunchKernelGGL(k_mlp_pack, dim3((kPackElems + 255) / 256), dim3(256), 0, c->stream, (const float*)m->P.p,
                       (__bf16*)m->pk.p);
    RRH_LAUNCH_CHECK();
    return RRH_OK;
}

template <class Src>
int mlp_grad(rrh_ctx* c, rrh_mlp* m, const Src& src, uint32_t n) {
    const size_t smem = 0;  // static LDS
    hipLaunchKernelGGL(k_mlp_train<Src>, dim3(m->nb), dim3(256), smem, c->stream, m->dev(), src, n, (const float*)m->wsum(),
                       (const float*)m->scale(), (float*)m->partial.p, (float*)m->ploss.p);
    RRH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_mlp_reduce, dim3((NPARAM + 255) / 256), dim3(256), 0, c->stream, (const float*)m->partial.p,
                       (const float*)m->ploss.p, m->nb, (const float*)m->scale(), (float*)m->grad.p, m->loss(), m->bad());
    RRH_LAUNCH_CHECK();
    return RRH_OK;
}

int mlp_update(rrh_ctx* c, rrh_mlp* m) {
    AdamArgs A{m->p.lr, m->p.beta1, m->p.beta2, m->p.eps};
    hipLaunchKernelGGL(k_mlp_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, c->stream, (float*)m->P.p,
                       (const float*)m->grad.p, (float*)m->m.p, (float*)m->v.p, (const uint32_t*)m->state(),
                       (const uint32_t*)m->bad(), A);
    RRH_LAUNCH_CHECK();
    int rc = mlp_pack(c, m);
    if (rc) return rc;
    hipLaunchKernelGGL(k_mlp_finish, dim3(1), dim3(64), 0, c->stream, m->state(), m->bad(), m->scale());
    RRH_LAUNCH_CHECK();
    return RRH_OK;
}

struct HostTmp {
    void* p[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    ~HostTmp() { for (void* q : p) if (q) hipFree(q); }
};
int upload_samples(rrh_ctx* c, const float* pos, const float* nrm, const float* wo, const float* dir, const float* w,
                   uint32_t n, HostTmp& T, HostSrc& src) {
    const float* h[5] = {pos, nrm, wo, dir, w};
    size_t bytes[5] = {(size_t)n * 12, (size_t)n * 12, (size_t)n * 12, (size_t)n * 12, (size_t)n * 4};
    for (int k = 0; k < 5; ++k) {
        if (!h[k]) continue;
        RRH_HIP(hipMalloc(&T.p[k], bytes[k]));
        RRH_HIP(hipMemcpyAsync(T.p[k], h[k], bytes[k], hipMemcpyHostToDevice, c->stream));
    }
    src = HostSrc{(const float*)T.p[0], (const float*)T.p[1], (const float*)T.p[2], (const float*)T.p[3],
                  (const float*)T.p[4], n};
    return RRH_OK;
}

}  // namespace

extern "C" {

int rrh_mlp_create(rrh_ctx* c, const rrh_mlp_params* p, rrh_mlp** out) {
    RRH_ARG(c && p && out, "rrh_mlp_create: null argument");
    for (int k = 0; k < 3; ++k)
        RRH_ARG(p->bounds_max[k] > p->bounds_min[k], "rrh_mlp_create: empty bounds");
    RRH_ARG(p->lr > 0.0f && p->beta1 >= 0.0f && p->beta1 < 1.0f && p->beta2 >= 0.0f && p->beta2 < 1.0f && p->eps > 0.0f,
            "rrh_mlp_create: bad Adam hyper-parameters");
    *out = nullptr;
    rrh_mlp* m = new rrh_mlp();
    m->ctx = c;
    m->p = *p;
    int cu = 256;
    hipDeviceProp_t pr;
    if (hipGetDeviceProperties(&pr, c->device) == hipSuccess) cu = pr.multiProcessorCount;
    m->nb = (uint32_t)cu;
    int rc = RRH_OK;
    DevBuf* bufs[] = {&m->P, &m->m, &m->v, &m->grad};
    for (DevBuf* b : bufs) if (!rc) rc = b->ensure((size_t)NPARAM * 4);
    if (!rc) rc = m->pk.ensure((size_t)kPackElems * 2);
    if (!rc) rc = m->partial.ensure((size_t)m->nb * NPARAM * 4);
    if (!rc) rc = m->ploss.ensure((size_t)m->nb * 4);
    if (!rc) rc = m->scal.ensure(64);
    if (rc) { rrh_mlp_destroy(m); return rc; }
    // Kaiming-uniform initialisation (bound sqrt(6 / fan_in)) from the seed; biases 0
    std::vector<float> P(NPARAM, 0.0f);
    uint32_t s = p->seed * 2654435761u + 12345u;
    for (int l = 0; l < 5; ++l) {
        float bound = std::sqrt(6.0f / (float)in_of(l));
        for (int k = 0; k < in_of(l) * out_of(l); ++k) {
            s = pcg_host(s);
            P[off_w(l) + k] = ((float)(s >> 8) * (1.0f / 16777216.0f) * 2.0f - 1.0f) * bound * (l == 4 ? 0.1f : 1.0f);
        }
    }
    float sc[8] = {0.0f, p->loss_scale > 0.0f ? p->loss_scale : 1024.0f, 0.0f, 0.0f};
    uint32_t st[2] = {0, 0};
    hipError_t e = hipMemcpyAsync(m->P.p, P.data(), (size_t)NPARAM * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(m->m.p, 0, (size_t)NPARAM * 4, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(m->v.p, 0, (size_t)NPARAM * 4, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(m->scal.p, 0, 64, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(m->scal.p, sc, 16, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(m->state(), st, 8, hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) { rrh_mlp_destroy(m); set_error(hipGetErrorString(e)); return RRH_E_HIP; }
    rc = mlp_pack(c, m);
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = RRH_E_HIP;
    if (rc) { rrh_mlp_destroy(m); return rc; }
    *out = m;
    return RRH_OK;
}

int rrh_mlp_destroy(rrh_mlp* m) {
    if (!m) return RRH_OK;
    (void)hipDeviceSynchronize();  // the owning context may already be gone
    DevBuf* bufs[] = {&m->P, &m->m, &m->v, &m->grad, &m->pk, &m->partial, &m->ploss, &m->scal};
    for (DevBuf* b : bufs) b->release();
    delete m;
    return RRH_OK;
}

int rrh_mlp_set_weights(rrh_ctx* c, rrh_mlp* m, const float* params, uint32_t n) {
    RRH_ARG(c && m && params, "rrh_mlp_set_weights: null argument");
    RRH_ARG(n == (uint32_t)NPARAM, "rrh_mlp_set_weights: expected RRH_MLP_PARAMS values");
    RRH_HIP(hipMemcpyAsync(m->P.p, params, (size_t)NPARAM * 4, hipMemcpyHostToDevice, c->stream));
    int rc = mlp_pack(c, m);
    if (rc) return rc;
    RRH_HIP(hipStreamSynchronize(c->stream));
    return RRH_OK;
}

int rrh_mlp_get_weights(rrh_ctx* c, const rrh_mlp* m, float* params, uint32_t n) {
    RRH_ARG(c && m && params, "rrh_mlp_get_weights: null argument");
    RRH_ARG(n == (uint32_t)NPARAM, "rrh_mlp_get_weights: expected RRH_MLP_PARAMS values");
    RRH_HIP(hipMemcpyAsync(params, m->P.p, (size_t)NPARAM * 4, hipMemcpyDeviceToHost, c->stream));
    RRH_HIP(hipStreamSynchronize(c->stream));
    return RRH_OK;
}

int rrh_mlp_infer(rrh_ctx* c, rrh_mlp* m, const float* pos, const float* nrm, const float* wo, uint32_t n,
                  float* lobes) {
    RRH_ARG(c && m && pos && nrm && wo && lobes, "rrh_mlp_infer: null argument");
    if (n == 0) return RRH_OK;
    HostTmp T;
    HostSrc src;
    int rc = upload_samples(c, pos, nrm, wo, nullptr, nullptr, n, T, src);
    if (rc) return rc;
    float* dl = nullptr;
    RRH_HIP(hipMalloc(&dl, (size_t)n * NK * 5 * 4));
    uint32_t g = (n + kBlockSamples - 1) / kBlockSamples;
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_mlp_infer<HostSrc>, dim3(g), dim3(256), 0, c->stream, m->dev(), src, n, dl);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(lobes, dl, (size_t)n * NK * 5 * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    hipFree(dl);
    if (e != hipSuccess) { set_error(std::string("rrh_mlp_infer: ") + hipGetErrorString(e)); return RRH_E_HIP; }
    return RRH_OK;
}

int rrh_mlp_loss_grad(rrh_ctx* c, rrh_mlp* m, const float* pos, const float* nrm, const float* wo, const float* dirs,
                      const float* w, uint32_t n, float* loss, float* grad) {
    RRH_ARG(c && m && pos && nrm && wo && dirs && w && loss && grad, "rrh_mlp_loss_grad: null argument");
    RRH_ARG(n > 0, "rrh_mlp_loss_grad: no samples");
    HostTmp T;
    HostSrc src;
    int rc = upload_samples(c, pos, nrm, wo, dirs, w, n, T, src);
    if (rc) return rc;
    RRH_HIP(hipMemsetAsync(m->wsum(), 0, 4, c->stream));
    hipLaunchKernelGGL(k_mlp_wsum_host, dim3(64), dim3(256), 0, c->stream, src.w, n, m->wsum());
    RRH_LAUNCH_CHECK();
    if ((rc = mlp_grad(c, m, src, n))) return rc;
    RRH_HIP(hipMemcpyAsync(grad, m->grad.p, (size_t)NPARAM * 4, hipMemcpyDeviceToHost, c->stream));
    RRH_HIP(hipMemcpyAsync(loss, m->loss(), 4, hipMemcpyDeviceToHost, c->stream));
    RRH_HIP(hipMemsetAsync(m->bad(), 0, 4, c->stream));
    RRH_HIP(hipStreamSynchronize(c->stream));
    return RRH_OK;
}

int rrh_mlp_train_step(rrh_ctx* c, rrh_mlp* m, const float* pos, const float* nrm, const float* wo, const float* dirs,
                       const float* w, uint32_t n, float* loss) {
    RRH_ARG(c && m && pos && nrm && wo && dirs && w, "rrh_mlp_train_step: null argument");
    RRH_ARG(n > 0, "rrh_mlp_train_step: no samples");
    HostTmp T;
    HostSrc src;
    int rc = upload_samples(c, pos, nrm, wo, dirs, w, n, T, src);
    if (rc) return rc;
    RRH_HIP(hipMemsetAsync(m->wsum(), 0, 4, c->stream));
    hipLaunchKernelGGL(k_mlp_wsum_host, dim3(64), dim3(256), 0, c->stream, src.w, n, m->wsum());
    RRH_LAUNCH_CHECK();
    if ((rc = mlp_grad(c, m, src, n))) return rc;
    if ((rc = mlp_update(c, m))) return rc;
    if (loss) RRH_HIP(hipMemcpyAsync(loss, m->loss(), 4, hipMemcpyDeviceToHost, c->stream));
    RRH_HIP(hipStreamSynchronize(c->stream));
    return RRH_OK;
}

int rrh_mlp_train_from_vertices(rrh_ctx* c, rrh_mlp* m, uint32_t batch, float* loss) {
    RRH_ARG(c && m, "rrh_mlp_train_from_vertices: null argument");
    RRH_ARG(c->train.p && c->train_cnt.p, "rrh_mlp_train_from_vertices: no recorded vertices (render with record_vertices)");
    RRH_ARG(batch > 0, "rrh_mlp_train_from_vertices: batch must be > 0");
    TrainSrc src{(const float4*)c->train.p, (const uint32_t*)c->train_cnt.p, c->train_capacity, batch};
    RRH_HIP(hipMemsetAsync(m->wsum(), 0, 4, c->stream));
    hipLaunchKernelGGL(k_mlp_wsum_train, dim3(64), dim3(256), 0, c->stream, src.tr, src.cnt, src.cap, batch, m->wsum());
    RRH_LAUNCH_CHECK();
    int rc = mlp_grad(c, m, src, batch);
    if (rc) return rc;
    if ((rc = mlp_update(c, m))) return rc;
    if (loss) {
        RRH_HIP(hipMemcpyAsync(loss, m->loss(), 4, hipMemcpyDeviceToHost, c->stream));
        RRH_HIP(hipStreamSynchronize(c->stream));
    }
    return RRH_OK;
}

int rrh_mlp_adam_apply(rrh_ctx* c, rrh_mlp* m, const float* grad, uint32_t n) {
    RRH_ARG(c && m && grad, "rrh_mlp_adam_apply: null argument");
    RRH_ARG(n == (uint32_t)NPARAM, "rrh_mlp_adam_apply: expected RRH_MLP_PARAMS values");
    RRH_HIP(hipMemcpyAsync(m->grad.p, grad, (size_t)NPARAM * 4, hipMemcpyHostToDevice, c->stream));
    int rc = mlp_update(c, m);
    if (rc) return rc;
    RRH_HIP(hipStreamSynchronize(c->stream));
    return RRH_OK;
}

int rrh_mlp_state(rrh_ctx* c, const rrh_mlp* m, uint32_t* step, float* loss_scale) {
    RRH_ARG(c && m && step && loss_scale, "rrh_mlp_state: null argument");
    RRH_HIP(hipMemcpyAsync(step, m->state(), 4, hipMemcpyDeviceToHost, c->stream));
    RRH_HIP(hipMemcpyAsync(loss_scale, m->scale(), 4, hipMemcpyDeviceToHost, c->stream));
    RRH_HIP(hipStreamSynchronize(c->stream));
    return RRH_OK;
}

}  // extern "C"

// ---- guided rendering: per-queue-slot lobes before shading --------------------------------
namespace rrh {

struct QueueSrc {  // hit points of the live ray queue (miss -> invalid)
    DevScene S;
    const float4* o;
    const float4* d;
    const float4* hits;
    const uint32_t* cnt;
    uint32_t b;
    __device__ uint32_t count() const { return cnt[b]; }
    __device__ bool get(uint32_t i, Sample& s) const {
        if (i >= cnt[b]) return false;
        float4 h = hits[i];
        uint32_t prim = __float_as_uint(h.y);
        if (prim == ~0u) return false;
        HitInfo hi = fetch_hit(S, prim, h.z, h.w);
        f3 dir = xyz(d[i]);
        f3 wo = neg(dir);
        f3 n = hi.normal;
        if (dot(n, wo) < 0.0f) n = neg(n);
        s.p = hi.pos; s.n = n; s.wo = wo; s.dir = mk(0, 0, 1); s.w = 0.0f;
        return true;
    }
};

int mlp_infer_queue(rrh_ctx* c, rrh_mlp* m, const DevScene& S, const float4* qo, const float4* qd, const float4* hits,
                    const uint32_t* cnt, uint32_t b, uint32_t cap, float* lobes) {
    QueueSrc src{S, qo, qd, hits, cnt, b};
    uint32_t g = (cap + kBlockSamples - 1) / kBlockSamples;
    if (g > 2048) g = 2048;
    hipLaunchKernelGGL(k_mlp_infer<QueueSrc>, dim3(g), dim3(256), 0, c->stream, m->dev(), src, cap, lobes);
    RRH_LAUNCH_CHECK();
    return RRH_OK;
}

}  // namespace rrh

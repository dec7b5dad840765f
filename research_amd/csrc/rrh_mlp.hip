// rrh_mlp.hip -- learned path guide: a small MLP that maps (position, normal, outgoing
// direction) to an 8-lobe vMF mixture, trained online on the path tracer's own vertices.
// The reference has no learned guide (its guide is the single-lobe grid of
// GuidingGrid.slang); the network and loss are defined in include/rrh.h.
//
// Training step = k_mlp_wsum (per-block sums of sample weights) -> k_mlp_train (persistent blocks:
// per 64-sample chunk, forward through five BF16 MFMA layers, vMF-mixture NLL head,
// backward deltas, then the block's weight-gradient tiles, kept in registers across
// chunks) -> k_mlp_reduce1/2 (fixed-order sum of the per-block partials, non-finite
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
constexpr int kWsumBlocks = 64;     // blocks of the sample-weight sum

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
// inference copy: the 42 MFMA A fragments of the five layers (layer 0: 4 x 3, layers 1-3: 4 x 2,
// layer 4: 3 x 2), each 64 lanes x 8 BF16 in lane order (1 KB, one ds_read_b128 per lane), layers
// 1..4 with permuted columns and layer 4 with permuted rows: see k_mlp_infer
constexpr int kFrags = 42;
constexpr int kFWOff = kWTOff + 4 * 4096;
constexpr int kPackElems = kFWOff + kFrags * 512;
__host__ __device__ constexpr int frag_index(int l, int nt, int kk) {
    return l == 0 ? nt * 3 + kk : (l < 4 ? 12 + (l - 1) * 8 + nt * 2 + kk : 36 + nt * 2 + kk);
}

struct MlpDev {
    const __bf16* pk;  // packed bf16 weights
    const float* P;    // fp32 master parameters (biases are read from here)
    float bmin[3], binv[3];  // position normalisation: u = (p - bmin) * binv
};

__device__ __forceinline__ const __bf16* Wl(const MlpDev& M, int l) { return M.pk + kWOff[l]; }
__device__ __forceinline__ const __bf16* WTl(const MlpDev& M, int l) { return M.pk + kWTOff + (l - 1) * 4096; }

// Inference runs each layer transposed, D = W X^T: a lane (r, g = lane / 16) of output tile nt
// holds features 16 nt + 4 g + j (j < 4) of sample r.  Tiles 2 kk and 2 kk + 1 together hold
// what the next layer's B operand needs at "virtual" columns 32 kk + 8 g + i, so the inference
// fragments of layers 1..4 hold column v = feature perm_k(v), and layer 4's rows are ordered so
// that lane g of sample r receives the raw outputs of lobes 2g and 2g + 1 (perm_out4).
__host__ __device__ constexpr int perm_k(int v) {
    return 32 * (v >> 5) + ((v & 7) < 4 ? 4 * ((v >> 3) & 3) + (v & 7) : 16 + 4 * ((v >> 3) & 3) + (v & 7) - 4);
}
__host__ __device__ constexpr int perm_out4(int o) {  // permuted row of layer 4 -> raw output, -1 = padding
    return (4 * (o >> 4) + (o & 3)) < 10 ? 10 * ((o >> 2) & 3) + 4 * (o >> 4) + (o & 3) : -1;
}

// ---- inputs --------------------------------------------------------------------------
struct Sample { f3 p, n, wo, dir; float w; };

struct HostSrc {  // arrays uploaded from the host (3 floats per vector)
    const float *pos, *nrm, *wo, *dir, *w;
    uint32_t n;
    __device__ uint32_t count() const { return n; }
    __device__ float weight(uint32_t i) const { return w[i]; }
    __device__ bool load(uint32_t i, Sample& s) const {  // i < n: the inference inputs only
        s.p = mk(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        s.n = mk(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
        s.wo = mk(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]);
        return true;
    }
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
// samples finalised from path-vertex records: t0 (pos, w_Li) t1 (dir, pdf) t2 (n, w_product) t3 (wo).
// The network is trained towards the product Li f cos (w_product), the ideal sampling density
// for the integrand at the vertex; the grid trains on incident radiance (w_Li).  When more
// than `batch` are available, the batch takes every (avail / batch)-th one, spread over the frame
struct TrainSrc {
    const float4* tr;
    const uint32_t* cnt;
    uint32_t cap, batch;
    __device__ uint32_t avail() const { return min(*cnt, cap); }
    __device__ uint32_t count() const { return min(avail(), batch); }
    __device__ uint32_t pick(uint32_t j, uint32_t a) const { return (a > batch) ? (uint32_t)(((uint64_t)j * a) / batch) : j; }
    __device__ float weight(uint32_t j) const { return tr[(size_t)pick(j, avail()) * 4 + 2].w; }
    __device__ bool get(uint32_t j, Sample& s) const {
        const uint32_t a = avail();
        if (j >= min(a, batch)) return false;
        const uint32_t i = pick(j, a);
        float4 t0 = tr[(size_t)i * 4], t1 = tr[(size_t)i * 4 + 1], t2 = tr[(size_t)i * 4 + 2], t3 = tr[(size_t)i * 4 + 3];
        s.p = xyz(t0); s.w = t2.w; s.dir = xyz(t1); s.n = xyz(t2); s.wo = xyz(t3);  // product-weighted target
        return true;
    }
};

// ---- encoding: features 0..71 = sin/cos(2^j pi u_c), 72..74 n, 75..77 wo, 78..95 zero ----
// A block of 8 features holds one coordinate at four consecutive frequencies j0..j0+3: the
// first pair comes from the hardware sine / cosine of the exactly reduced angle (argument in
// revolutions, (2^j0 u mod 2) / 2), the next three by angle doubling (sin 2x = 2 s c,
// cos 2x = 1 - 2 s^2); the error (~1e-5 after three doublings) is far below the BF16
// rounding each feature gets next.
__device__ __forceinline__ float unit_coord(const MlpDev& M, const Sample& s, int c) {
    const float pc = (c == 0) ? s.p.x : (c == 1 ? s.p.y : s.p.z);
    const float lo = (c == 0) ? M.bmin[0] : (c == 1 ? M.bmin[1] : M.bmin[2]);  // selects: no indexed
    const float inv = (c == 0) ? M.binv[0] : (c == 1 ? M.binv[1] : M.binv[2]); // copy of the argument
    return clampf((pc - lo) * inv, 0.0f, 1.0f);
}
__device__ __forceinline__ void freq_pair(float u, int j, float& sn, float& cs) {
    const float t = u * (float)(1 << j);           // exact scaling by 2^j
    const float rev = 0.5f * (t - 2.0f * floorf(t * 0.5f));  // (t mod 2) / 2, exact for t < 2^24
    sn = __builtin_amdgcn_sinf(rev);
    cs = __builtin_amdgcn_cosf(rev);
}
// two features rounded to BF16 by one paired conversion
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t bf16x2_bits(float a, float b) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){a, b}, bf16x2));
}
// features [fb, fb + 8) of a sample (fb a multiple of 8); MASK false leaves an invalid sample's
// features as they come (inference: a sample's column of the MFMA operands reaches only its own
// outputs, and an invalid sample's outputs are never stored)
template <bool MASK = true>
__device__ __forceinline__ bf16x8 encode8(const MlpDev& M, const Sample& s, bool valid, int fb) {
    uint32_t q[4];
    if (fb < 72) {
        float sn, cs;
        freq_pair(unit_coord(M, s, fb / 24), (fb % 24) >> 1, sn, cs);
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            q[p] = bf16x2_bits(!MASK || valid ? sn : 0.0f, !MASK || valid ? cs : 0.0f);
            const float t2 = sn + sn;  // sin 2x = (2 s) c, cos 2x = 1 - (2 s) s
            const float s2 = t2 * cs, c2 = __builtin_fmaf(-t2, sn, 1.0f);
            sn = s2;
            cs = c2;
        }
    } else {
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            const int k = fb - 72 + 2 * p;  // 0..23: n (3), wo (3), zero padding
            const float v0 = (k == 0) ? s.n.x : (k == 2) ? s.n.z : (k == 4) ? s.wo.y : 0.0f;
            const float v1 = (k == 0) ? s.n.y : (k == 2) ? s.wo.x : (k == 4) ? s.wo.z : 0.0f;
            q[p] = bf16x2_bits(!MASK || valid ? v0 : 0.0f, !MASK || valid ? v1 : 0.0f);
        }
    }
    return __builtin_bit_cast(bf16x8, (uint4){q[0], q[1], q[2], q[3]});
}
// 4 lanes per sample; lane quarter q writes features [24q, 24q + 24) into the sample's row
__device__ __forceinline__ void encode_row(const MlpDev& M, const Sample& s, bool valid, int q, __bf16* row) {
#pragma unroll
    for (int k = 0; k < 3; ++k) *reinterpret_cast<bf16x8*>(row + q * 24 + 8 * k) = encode8(M, s, valid, q * 24 + 8 * k);
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
// The head uses the hardware exponential and logarithm (~1 ulp): the learned guide is held
// to tolerances, not to bit parity with the FP64 oracle.
__device__ __forceinline__ float hexp(float x) { return __expf(x); }
__device__ __forceinline__ float hlog(float x) { return __logf(x); }
__device__ __forceinline__ float softplus_f(float d) { return d > 20.0f ? d : hlog(1.0f + hexp(d)); }
__device__ __forceinline__ float sigmoid_f(float d) { return 1.0f / (1.0f + hexp(-d)); }

// The head runs on all 64 lanes: the 4 adjacent lanes of a sample each own two lobes
// (lane quarter qd: lobes 2 qd, 2 qd + 1), so the softmax and mixture sums are two xor shuffles.
struct Lobe { float mu[3], kap, pi; };
__device__ __forceinline__ float group_sum4(float v) {
    v += __shfl_xor(v, 1);
    return v + __shfl_xor(v, 2);
}
__device__ __forceinline__ float group_max4(float v) {
    v = fmaxf(v, __shfl_xor(v, 1));
    return fmaxf(v, __shfl_xor(v, 2));
}
__device__ __forceinline__ void head_lobes2(const float* o, int qd, Lobe* lb) {
    const float e0 = o[(2 * qd) * 5 + 4], e1 = o[(2 * qd + 1) * 5 + 4];
    const float emax = group_max4(fmaxf(e0, e1));
    const float x[2] = {hexp(e0 - emax), hexp(e1 - emax)};
    const float z = group_sum4(x[0] + x[1]);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const float* v = o + (2 * qd + j) * 5;
        float len = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + 1e-12f);
        lb[j].mu[0] = v[0] / len;
        lb[j].mu[1] = v[1] / len;
        lb[j].mu[2] = v[2] / len;
        lb[j].kap = fminf(0.1f + softplus_f(v[3]), 1000.0f);
        lb[j].pi = x[j] / z;
    }
}
__device__ __forceinline__ float vmf_val(float kap, float c) {
    return kap * hexp(kap * (c - 1.0f)) / (6.28318530717958647692f * (1.0f - hexp(-2.0f * kap)));
}
// NLL head of one sample on its 4 lanes: returns the loss term (on lane quarter 0, else 0)
// and writes dL/dout (times `gscale`) of this lane's two lobes into g[10]
__device__ __forceinline__ float head_grad2(const float* o, int qd, f3 d, float wn, float gscale, float* g) {
    Lobe lb[2];
    head_lobes2(o, qd, lb);
    float c[2], f[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        c[j] = lb[j].mu[0] * d.x + lb[j].mu[1] * d.y + lb[j].mu[2] * d.z;
        f[j] = vmf_val(lb[j].kap, c[j]);
    }
    const float q = group_sum4(lb[0].pi * f[0] + lb[1].pi * f[1]);
    const bool on = wn > 0.0f;
    const float dq = on ? -wn / (q + 1e-20f) * gscale : 0.0f;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const float* v = o + (2 * qd + j) * 5;
        const float pik = lb[j].pi, kap = lb[j].kap;
        g[j * 5 + 4] = dq * pik * (f[j] - q);
        float e2 = hexp(-2.0f * kap);
        float dlogf = 1.0f / kap - 2.0f * e2 / (1.0f - e2) + (c[j] - 1.0f);
        float raw = 0.1f + softplus_f(v[3]);
        g[j * 5 + 3] = (raw > 1000.0f) ? 0.0f : dq * pik * f[j] * dlogf * sigmoid_f(v[3]);
        float len = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + 1e-12f);
        float s = dq * pik * f[j] * kap;
        float g0 = s * d.x, g1 = s * d.y, g2 = s * d.z;
        float mg = lb[j].mu[0] * g0 + lb[j].mu[1] * g1 + lb[j].mu[2] * g2;
        g[j * 5 + 0] = (g0 - lb[j].mu[0] * mg) / len;
        g[j * 5 + 1] = (g1 - lb[j].mu[1] * mg) / len;
        g[j * 5 + 2] = (g2 - lb[j].mu[2] * mg) / len;
    }
    return (on && qd == 0) ? -wn * hlog(q + 1e-20f) : 0.0f;
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
        mine = s;  // every lane of the sample keeps it (the head runs on all four)
        valid = v;
    }
    __syncthreads();
    layer_fwd<KI, 4>(X + w * 16 * KI, Wl(M, 0), M.P + off_b(0), NH, true, H[0] + w * 16 * 64, nullptr);
    __syncthreads();
#pragma unroll
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
    float wt = 0.0f;  // total sample weight: the kWsumBlocks partials, summed in order
    for (int b = 0; b < kWsumBlocks; ++b) wt += wsum[b];
    const float inv_w = 1.0f / fmaxf(wt, 1e-30f);
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
        // head: 4 lanes per sample
        {
            const int sl = lane >> 2, qd = lane & 3;
            const float wn = valid ? mine.w * inv_w : 0.0f;
            float g[10];
            loss_acc += head_grad2(sm.O(w) + sl * NOP, qd, mine.dir, wn, gscale, g);
            __bf16* Drow = sm.D[4] + (w * 16 + sl) * 64;
#pragma unroll
            for (int k = 0; k < 10; ++k) Drow[qd * 10 + k] = (__bf16)g[k];
#pragma unroll
            for (int k = 0; k < 6; ++k) Drow[NO + qd * 6 + k] = (__bf16)0.0f;
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

// Inference: persistent waves.  The block stages the 42 weight fragments (42 KB, fragment-major:
// one conflict-free ds_read_b128 per lane and fragment) and the biases (layer 4's in permuted
// order) in LDS once; each wave then takes 32 samples per step as two 16-sample tiles, and every
// fragment read feeds both tiles' MFMAs.  The layers run transposed (D = W X^T, see perm_k):
// each layer's output fragments become the next layer's input fragments after bias, ReLU and
// BF16 rounding, in registers; layer-0 operands are encoded straight into registers and the head
// runs on the four lanes (r, r + 16, r + 32, r + 48) that hold a sample's lobes.
// ReLU after the BF16 rounding, as a packed 16-bit integer max with zero (one instruction per
// pair): round-to-nearest keeps the sign, and a negative BF16 (or -0) is negative as an integer,
// so this equals rounding the ReLU of the FP32 value
typedef short i16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t relu_bf16x2(float a, float b) {  // one v_cvt_pk_bf16_f32 + v_pk_max_i16
    const bf16x2 h = __builtin_convertvector((f32x2){a, b}, bf16x2);
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(i16x2, h), (i16x2)0));
}
__device__ __forceinline__ bf16x8 relu_bf16x8(const f32x4& lo, const f32x4& hi) {
    const uint4 q = {relu_bf16x2(lo[0], lo[1]), relu_bf16x2(lo[2], lo[3]), relu_bf16x2(hi[0], hi[1]),
                     relu_bf16x2(hi[2], hi[3])};
    return __builtin_bit_cast(bf16x8, q);
}
__device__ __forceinline__ float group_sum4t(float v) {
    v += __shfl_xor(v, 16);
    return v + __shfl_xor(v, 32);
}
__device__ __forceinline__ float group_max4t(float v) {
    v = fmaxf(v, __shfl_xor(v, 16));
    return fmaxf(v, __shfl_xor(v, 32));
}
__device__ __forceinline__ bf16x8 frag(const bf16x8* s_w, int f) { return s_w[f * 64 + (threadIdx.x & 63)]; }

// hidden layer l (1..3) of both tiles: h[t] in, h[t] out
__device__ __forceinline__ void hidden2(const bf16x8* s_w, const float* bias, int l, bf16x8 (&h)[2][2]) {
    const int g = (threadIdx.x & 63) >> 4;
    f32x4 d[2][4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        const f32x4 b = *reinterpret_cast<const f32x4*>(bias + nt * 16 + 4 * g);
        d[0][nt] = b;
        d[1][nt] = b;
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
            const bf16x8 a = frag(s_w, frag_index(l, nt, kk));
            d[0][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, h[0][kk], d[0][nt], 0, 0, 0);
            d[1][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, h[1][kk], d[1][nt], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        h[t][0] = relu_bf16x8(d[t][0], d[t][1]);
        h[t][1] = relu_bf16x8(d[t][2], d[t][3]);
    }
}
// head of one tile: this lane holds the raw outputs of lobes 2g (o[0..4]) and 2g + 1 (o[5..9])
// of sample `si`; when the sample is valid, writes them as (mu3, kappa, pi) floats (H16 false:
// 40 floats per sample) or as the fp16 record LobesH16 reads (8 x (mu3, kappa), then
// the 8 weights)
template <bool H16>
__device__ __forceinline__ void head_store_t(const f32x4 (&e)[3], bool valid, void* out, size_t si) {
    const int g = (threadIdx.x & 63) >> 4;
    float o[10];
#pragma unroll
    for (int q = 0; q < 10; ++q) o[q] = e[q >> 2][q & 3];
    const float emax = group_max4t(fmaxf(o[4], o[9]));
    const float x0 = hexp(o[4] - emax), x1 = hexp(o[9] - emax);
    const float z = group_sum4t(x0 + x1);
    if (!valid) return;
    // hardware reciprocal square root and reciprocal (about 1 ulp) in place of IEEE divisions
    const float rz = __builtin_amdgcn_rcpf(z);
    float lb[10];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const float* v = o + 5 * j;
        const float il = __builtin_amdgcn_rsqf(__builtin_fmaf(v[0], v[0], __builtin_fmaf(v[1], v[1], v[2] * v[2])) + 1e-12f);
        lb[j * 5 + 0] = v[0] * il;
        lb[j * 5 + 1] = v[1] * il;
        lb[j * 5 + 2] = v[2] * il;
        // softplus with the bare base-2 logarithm: its argument 1 + e^d is at least 1
        const float sp = v[3] > 20.0f ? v[3] : 0.69314718f * __builtin_amdgcn_logf(1.0f + hexp(v[3]));
        lb[j * 5 + 3] = fminf(0.1f + sp, 1000.0f);
        lb[j * 5 + 4] = (j == 0 ? x0 : x1) * rz;
    }
    if constexpr (H16) {
        typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
        uint16_t* rec = reinterpret_cast<uint16_t*>(out) + si * kMlpLobeHalves;
        auto h2 = [](float a, float b) {  // one v_cvt_pk_f16_f32 (round to nearest even)
            return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){a, b}, f16x2));
        };
        uint4 q;
        q.x = h2(lb[0], lb[1]);
        q.y = h2(lb[2], lb[3]);
        q.z = h2(lb[5], lb[6]);
        q.w = h2(lb[7], lb[8]);
        *reinterpret_cast<uint4*>(rec + 8 * g) = q;
        *reinterpret_cast<uint32_t*>(rec + 32 + 2 * g) = h2(lb[4], lb[9]);
    } else {
        float* dst = reinterpret_cast<float*>(out) + si * (NK * 5) + g * 10;
#pragma unroll
        for (int q = 0; q < 10; ++q) dst[q] = lb[q];
    }
}

// 8 waves per block share one copy of the fragments (44 KB of LDS): three blocks per CU, and the
// registers (about 110 per lane) allow four waves per SIMD
constexpr int kInferWaves = 8;
template <class Src, bool H16>
__global__ void __launch_bounds__(64 * kInferWaves, 4) k_mlp_infer(MlpDev M, Src src, uint32_t nsamples,
                                                               void* __restrict__ lobes) {
    __shared__ bf16x8 s_w[kFrags * 64];
    __shared__ __align__(16) float s_bias[4 * 64 + 48];
    const uint4* gw = reinterpret_cast<const uint4*>(M.pk + kFWOff);
    for (int k = threadIdx.x; k < kFrags * 64; k += 64 * kInferWaves) reinterpret_cast<uint4*>(s_w)[k] = gw[k];
    for (int k = threadIdx.x; k < 4 * 64 + 48; k += 64 * kInferWaves) {
        float v;
        if (k < 256) {
            v = M.P[off_b(k >> 6) + (k & 63)];
        } else {
            const int raw = perm_out4(k - 256);
            v = raw >= 0 ? M.P[off_b(4) + raw] : 0.0f;
        }
        s_bias[k] = v;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    const uint32_t n = min(src.count(), nsamples);
    const uint32_t nchunks = (n + 31) / 32;
    const uint32_t stride = gridDim.x * kInferWaves;
    uint32_t ch = blockIdx.x * kInferWaves + (threadIdx.x >> 6);
    if (n == 0) return;
    // the next step's samples are loaded while this one computes (software pipelining); lanes
    // (r, g = 0..3) of tile t hold sample 32 ch + 16 t + r.  Loads are unconditional, at the index
    // clamped to the last sample (no branches, no register copies on the merge): a slot past the
    // end or a miss holds some sample's finite inputs, and its outputs are dropped
    auto fetch = [&](uint32_t i, Sample& s) { return src.load(min(i, n - 1), s) && i < n; };
    Sample nxt[2];
    bool nv[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) nv[t] = fetch(ch * 32 + t * 16 + r, nxt[t]);
    for (; ch < nchunks; ch += stride) {
        // the weight fragments stay in LDS: keep the compiler from hoisting their reads into
        // registers across steps (that would cost the occupancy this layout is for)
        asm volatile("" ::: "memory");
        const uint32_t base = ch * 32;
        Sample smp[2] = {nxt[0], nxt[1]};
        const bool valid[2] = {nv[0], nv[1]};
        const bool more = ch + stride < nchunks;
#pragma unroll
        for (int t = 0; t < 2; ++t) nv[t] = fetch((ch + stride) * 32 + t * 16 + r, nxt[t]) && more;
        bf16x8 h[2][2];
        {   // layer 0: B operand = features 32 kk + 8 g .. of sample r, encoded in registers
            bf16x8 x[2][3];
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int kk = 0; kk < 3; ++kk) x[t][kk] = encode8<false>(M, smp[t], valid[t], kk * 32 + g * 8);
            f32x4 d[2][4];
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                const f32x4 b = *reinterpret_cast<const f32x4*>(s_bias + nt * 16 + 4 * g);
                d[0][nt] = b;
                d[1][nt] = b;
#pragma unroll
                for (int kk = 0; kk < 3; ++kk) {
                    const bf16x8 a = frag(s_w, frag_index(0, nt, kk));
                    d[0][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, x[0][kk], d[0][nt], 0, 0, 0);
                    d[1][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, x[1][kk], d[1][nt], 0, 0, 0);
                }
            }
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                h[t][0] = relu_bf16x8(d[t][0], d[t][1]);
                h[t][1] = relu_bf16x8(d[t][2], d[t][3]);
            }
        }
        hidden2(s_w, s_bias + 64, 1, h);
        hidden2(s_w, s_bias + 128, 2, h);
        hidden2(s_w, s_bias + 192, 3, h);
        f32x4 e[2][3];
#pragma unroll
        for (int nt = 0; nt < 3; ++nt) {
            const f32x4 b = *reinterpret_cast<const f32x4*>(s_bias + 256 + nt * 16 + 4 * g);
            e[0][nt] = b;
            e[1][nt] = b;
#pragma unroll
            for (int kk = 0; kk < 2; ++kk) {
                const bf16x8 a = frag(s_w, frag_index(4, nt, kk));
                e[0][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, h[0][kk], e[0][nt], 0, 0, 0);
                e[1][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, h[1][kk], e[1][nt], 0, 0, 0);
            }
        }
#pragma unroll
        for (int t = 0; t < 2; ++t)
            head_store_t<H16>(e[t], valid[t], lobes, (size_t)(base + t * 16 + r));
    }
}

// per-block sums of the sample weights (fixed grid and order: deterministic)
template <class Src>
__global__ void __launch_bounds__(256) k_mlp_wsum(Src src, float* __restrict__ out) {
    __shared__ float red[4];
    const uint32_t n = src.count();
    float v = 0.0f;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v += src.weight(i);
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// two-stage fixed-order sum of the per-block partial gradients (deterministic):
// stage 1 sums groups of kGroup blocks, stage 2 the groups, unscales and flags non-finite values
constexpr uint32_t kGroup = 16;
__global__ void k_mlp_reduce1(const float* __restrict__ partial, uint32_t nb, float* __restrict__ tmp) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x, gq = blockIdx.y;
    if (i >= (uint32_t)NPARAM) return;
    const uint32_t b0 = gq * kGroup, b1 = min(b0 + kGroup, nb);
    float s = 0.0f;
    for (uint32_t b = b0; b < b1; ++b) s += partial[(size_t)b * NPARAM + i];
    tmp[(size_t)gq * NPARAM + i] = s;
}
__global__ void k_mlp_reduce2(const float* __restrict__ tmp, uint32_t ng, const float* __restrict__ partial_loss,
                              uint32_t nb, const float* __restrict__ scale_p, float* __restrict__ grad,
                              float* __restrict__ loss, uint32_t* __restrict__ bad) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    float inv = 1.0f / *scale_p;
    if (i < (uint32_t)NPARAM) {
        float s = 0.0f;
        for (uint32_t g = 0; g < ng; ++g) s += tmp[(size_t)g * NPARAM + i];
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

// bf16 copies: W_l [Opad][Kpad], WT_l [64][64] (l >= 1) and the inference fragments, zero padded
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
    } else if (e < kFWOff) {
        int local = e - kWTOff;
        int l = 1 + local / 4096;
        int i = (local % 4096) / 64, o = local % 64;
        if (o < out_of(l) && i < in_of(l)) val = P[off_w(l) + o * in_of(l) + i];
    } else {
        const int local = e - kFWOff;
        const int f = local / 512, lane = (local % 512) / 8, i = local % 8;
        const int l = f < 12 ? 0 : (f < 36 ? 1 + (f - 12) / 8 : 4);
        const int fl = f < 12 ? f : (f < 36 ? (f - 12) % 8 : f - 36);
        const int nt = l == 0 ? fl / 3 : fl / 2, kk = l == 0 ? fl % 3 : fl % 2;
        const int row = nt * 16 + (lane & 15), col = kk * 32 + (lane >> 4) * 8 + i;
        if (l == 0) {
            if (row < NH && col < NI) val = P[off_w(0) + row * NI + col];
        } else {
            const int o = (l == 4) ? perm_out4(row) : row;
            if (o >= 0 && o < out_of(l)) val = P[off_w(l) + o * in_of(l) + perm_k(col)];
        }
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
    DevBuf P, m, v, grad, pk, partial, part2, ploss, scal;
    // scal (512 B): float [1] loss scale, [2] loss; u32 [4] Adam steps, [5] good steps, [6] overflow flag;
    // float [16, 16 + kWsumBlocks) per-block sample-weight sums
    uint32_t nb = 0, cu_count = 256, infer_resident = 0;
    MlpDev dev() const {
        MlpDev d;
        d.pk = (const __bf16*)pk.p;
        d.P = (const float*)P.p;
        for (int k = 0; k < 3; ++k) { d.bmin[k] = p.bounds_min[k]; d.binv[k] = 1.0f / (p.bounds_max[k] - p.bounds_min[k]); }
        return d;
    }
    float* wsum() const { return (float*)scal.p + 16; }
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
    hipLaunchKernelGGL(k_mlp_wsum<Src>, dim3(kWsumBlocks), dim3(256), 0, c->stream, src, m->wsum());
    RRH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_mlp_train<Src>, dim3(m->nb), dim3(256), 0, c->stream, m->dev(), src, n, (const float*)m->wsum(),
                       (const float*)m->scale(), (float*)m->partial.p, (float*)m->ploss.p);
    RRH_LAUNCH_CHECK();
    const uint32_t ng = (m->nb + kGroup - 1) / kGroup;
    hipLaunchKernelGGL(k_mlp_reduce1, dim3((NPARAM + 255) / 256, ng), dim3(256), 0, c->stream,
                       (const float*)m->partial.p, m->nb, (float*)m->part2.p);
    RRH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_mlp_reduce2, dim3((NPARAM + 255) / 256), dim3(256), 0, c->stream, (const float*)m->part2.p,
                       ng, (const float*)m->ploss.p, m->nb, (const float*)m->scale(), (float*)m->grad.p, m->loss(),
                       m->bad());
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

// persistent inference grid: at most the blocks the device holds at once, fewer for small n
template <class K>
uint32_t infer_grid(rrh_ctx* c, rrh_mlp* m, K kernel, uint32_t n) {
    if (m->infer_resident == 0) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 64 * kInferWaves, 0) != hipSuccess || per_cu < 1)
            per_cu = 1;
        m->infer_resident = (uint32_t)per_cu * m->cu_count;
    }
    (void)c;
    const uint32_t need = (n + kInferWaves * 32 - 1) / (kInferWaves * 32);  // waves of 32-sample steps
    return need < 1 ? 1 : (need < m->infer_resident ? need : m->infer_resident);
}

struct HostTmp {
    void* p[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    ~HostTmp() { for (void* q : p) if (q) (void)hipFree(q); }
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
    m->cu_count = (uint32_t)cu;
    int rc = RRH_OK;
    DevBuf* bufs[] = {&m->P, &m->m, &m->v, &m->grad};
    for (DevBuf* b : bufs) if (!rc) rc = b->ensure((size_t)NPARAM * 4);
    if (!rc) rc = m->pk.ensure((size_t)kPackElems * 2);
    if (!rc) rc = m->partial.ensure((size_t)m->nb * NPARAM * 4);
    if (!rc) rc = m->part2.ensure((size_t)((m->nb + kGroup - 1) / kGroup) * NPARAM * 4);
    if (!rc) rc = m->ploss.ensure((size_t)m->nb * 4);
    if (!rc) rc = m->scal.ensure(512);
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
    if (e == hipSuccess) e = hipMemsetAsync(m->scal.p, 0, 512, c->stream);
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
    DevBuf* bufs[] = {&m->P, &m->m, &m->v, &m->grad, &m->pk, &m->partial, &m->part2, &m->ploss, &m->scal};
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
    const uint32_t g = infer_grid(c, m, k_mlp_infer<HostSrc, false>, n);
    hipLaunchKernelGGL((k_mlp_infer<HostSrc, false>), dim3(g), dim3(64 * kInferWaves), 0, c->stream, m->dev(), src, n,
                       (void*)dl);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(lobes, dl, (size_t)n * NK * 5 * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dl);
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

// The network's inputs for the live ray queue: hit point, normal facing the ray origin and
// outgoing direction (as the shade kernel derives them), written by a light, high-occupancy
// kernel so the persistent inference waves read them with one coalesced load instead of
// the chain of dependent fetches a hit lookup takes.  The record also carries the material
// index, and the shade kernel of the same bounce reads it in place of its own hit lookup.
__global__ void __launch_bounds__(256) k_mlp_prep(DevScene S, const float4* __restrict__ d,
                                                  const float4* __restrict__ hits, const uint32_t* __restrict__ cnt,
                                                  uint32_t b, float4* __restrict__ prep) {
    const uint32_t n = cnt[b];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        write_hit_prep(S, d[i], hits[i], prep, i);
}

struct QueueSrc {  // the prepared inputs of the live ray queue (miss -> invalid)
    const float4* prep;
    const uint32_t* cnt;
    uint32_t b;
    __device__ uint32_t count() const { return cnt[b]; }
    __device__ bool load(uint32_t i, Sample& s) const {  // i < count(); false for a miss
        const float4 p0 = prep[(size_t)i * 3], p1 = prep[(size_t)i * 3 + 1], p2 = prep[(size_t)i * 3 + 2];
        s.p = xyz(p0); s.n = xyz(p1); s.wo = xyz(p2);
        return p0.w > 0.0f;
    }
};

int mlp_infer_queue(rrh_ctx* c, rrh_mlp* m, const DevScene& S, const float4* qo, const float4* qd, const float4* hits,
                    const uint32_t* cnt, uint32_t b, uint32_t cap, uint16_t* lobes, bool prep_done) {
    (void)qo;
    int rc = c->mlp_prep.ensure((size_t)cap * 48);
    if (rc) return rc;
    if (!prep_done) {
        uint32_t gp = (cap + 255) / 256;
        gp = gp < 1u ? 1u : (gp > 4096u ? 4096u : gp);
        hipLaunchKernelGGL(k_mlp_prep, dim3(gp), dim3(256), 0, c->stream, S, qd, hits, cnt, b, (float4*)c->mlp_prep.p);
        RRH_LAUNCH_CHECK();
    }
    QueueSrc src{(const float4*)c->mlp_prep.p, cnt, b};
    const uint32_t g = infer_grid(c, m, k_mlp_infer<QueueSrc, true>, cap);
    hipLaunchKernelGGL((k_mlp_infer<QueueSrc, true>), dim3(g), dim3(64 * kInferWaves), 0, c->stream, m->dev(), src, cap,
                       (void*)lobes);
    RRH_LAUNCH_CHECK();
    return RRH_OK;
}

}  // namespace rrh

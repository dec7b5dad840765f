// rrh_guide.hip -- vMF path guiding.
//
// RRH_GRID_HASH restates the reference's ReSTIR-PT guiding grid:
//   cell index   GuidingGrid.slang:54-67     splat  :73-90 (int fixed-point atomics)
//   fit          :100-137 / cs_fit_guiding restir_pt.slang:1005-1015 (one thread per cell)
//   sample       sample_primary_dir restir_pt.slang:520-563
//   buffers      ReSTIRPTPass.cpp:311-341 (accum + fitted, 16 B per cell each)
// RRH_GRID_VMM is new: a res^3 dense grid with K <= 8 vMF lobes per cell fitted
// by online EM.  E-step: responsibilities in fp32, sufficient statistics
// (w*g, w*g*dir) in FP64.  The samples are sorted by cell first (rocPRIM radix sort
// of cell keys), then one wave per cell sums its samples in registers and one lane
// adds the 8 x 4 totals to the cell's frame statistics: no atomics, and the sums do
// not depend on scheduling.  M-step per cell: weights from
// the responsibility mass and the closed-form mean-resultant fit of the
// reference (GuidingGrid.slang:125-136) per lobe, so K = 1 reproduces it.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "rrh_internal.h"
#include "rrh_guide_dev.h"

using namespace rrh;

namespace {

__global__ void k_hash_splat(const float* __restrict__ pos, const uint32_t* __restrict__ bsdf,
                             const float* __restrict__ dir, uint32_t n, f3 cam, float base, float slope,
                             uint32_t count, uint32_t* __restrict__ accum) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    f3 p = mk(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    hash_splat(accum, hash_cell(p, bsdf[i], cam, base, slope, count), mk(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]));
}

__global__ void k_hash_fit(uint4* __restrict__ accum, float4* __restrict__ fitted, uint32_t count,
                           uint32_t min_samples, float kappa_max) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= count) return;
    uint4 a = accum[c];
    accum[c] = make_uint4(0, 0, 0, 0);
    uint32_t n = a.w;
    if (n < (min_samples > 1u ? min_samples : 1u)) {
        fitted[c].w = -1.0f;
        return;
    }
    f3 sum = scale(mk((float)(int32_t)a.x, (float)(int32_t)a.y, (float)(int32_t)a.z), 1.0f / 1048576.0f);
    float len = length(sum);
    float r = len / (float)n;
    r = clampf(r, 1.0e-4f, 0.9999f);
    f3 mu = (len > 1.0e-6f) ? divs(sum, len) : mk(0.0f, 0.0f, 1.0f);
    float kappa = r * (3.0f - r * r) / (1.0f - r * r);
    kappa = clampf(kappa, 0.5f, kappa_max);
    fitted[c] = make_float4(mu.x, mu.y, mu.z, kappa);
}

// restir_pt.slang:520-563
__global__ void k_hash_sample(const float4* __restrict__ fitted, const float* __restrict__ pos,
                              const float* __restrict__ nrm, const uint32_t* __restrict__ bsdf,
                              uint32_t* __restrict__ seeds, uint32_t n, f3 cam, float base, float slope,
                              uint32_t count, float alpha_in, float* __restrict__ dirs, float* __restrict__ pdfs) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    f3 p = mk(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    f3 N = mk(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
    uint32_t seed = seeds[i];
    const HashGrid G{fitted, nullptr, base, slope, count};
    const f3 wi = hash_primary_dir(G, p, N, bsdf[i], cam, alpha_in, seed, pdfs[i]);
    dirs[3 * i] = wi.x; dirs[3 * i + 1] = wi.y; dirs[3 * i + 2] = wi.z;
    seeds[i] = seed;
}

// ---- VMM ----------------------------------------------------------------------
struct VmmArgs {
    uint32_t res, K;
    float bmin[3], bext[3];
};

// E-step sample sources: host arrays uploaded (pos, dir, w: 3 + 3 + 1 floats) or the
// path vertices a guided render recorded (t0 = (pos, w), t1 = (dir, pdf); count on the device)
struct ArraySrc {
    const float *pos, *dir, *w;
    __device__ uint32_t count(uint32_t cap) const { return cap; }
    __device__ void get(uint32_t i, f3& p, f3& d, float& wt) const {
        p = mk(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        d = mk(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]);
        wt = w[i];
    }
};
struct TrainSrc {
    const float4* tr;
    const uint32_t* cnt;
    __device__ uint32_t count(uint32_t cap) const { return min(*cnt, cap); }
    __device__ void get(uint32_t i, f3& p, f3& d, float& wt) const {
        const float4 t0 = tr[(size_t)i * 4], t1 = tr[(size_t)i * 4 + 1];
        p = xyz(t0); d = xyz(t1); wt = t0.w;
    }
};

__device__ __forceinline__ uint32_t vmm_cell_of(const VmmArgs& A, f3 p) {
    GuideDev G;
    G.res = A.res; G.K = A.K;
    for (int k = 0; k < 3; ++k) { G.bmin[k] = A.bmin[k]; G.bext[k] = A.bext[k]; }
    return vmm_cell(G, p);
}

// key = the sample's cell; samples past the count or with no weight get the key ncell and
// sort behind every cell
template <class Src>
__global__ void k_em_keys(VmmArgs A, Src src, uint32_t cap, uint32_t ncell, uint32_t* __restrict__ keys,
                          uint32_t* __restrict__ vals) {
    const uint32_t n = src.count(cap);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += gridDim.x * blockDim.x) {
        uint32_t key = ncell;
        if (i < n) {
            f3 p, d;
            float w;
            src.get(i, p, d, w);
            if (w > 0.0f) key = vmm_cell_of(A, p);
        }
        keys[i] = key;
        vals[i] = i;
    }
}

// start[c] = first sorted position whose key is >= c, for c in [0, ncell]; every entry is
// written exactly once (the boundaries between consecutive keys cover [0, ncell])
__global__ void k_em_ranges(const uint32_t* __restrict__ keys, uint32_t n, uint32_t ncell, uint32_t* __restrict__ start) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += gridDim.x * blockDim.x) {
        const uint32_t lo = (i == 0) ? 0u : keys[i - 1] + 1u;  // cells in (keys[i-1], keys[i]] start at i
        const uint32_t hi = (i == n) ? ncell : keys[i];
        for (uint32_t c = lo; c <= hi; ++c) start[c] = i;
    }
}

// one wave per cell (grid-stride over cells): lanes stride over the cell's sorted samples,
// keep the K x 4 sums in registers, then an xor-shuffle tree reduces them and lane 0 adds
// them to the cell's frame statistics (cleared by the M-step)
template <class Src>
__global__ void __launch_bounds__(256) k_em_cells(VmmArgs A, Src src, const float* __restrict__ lobes,
                                                  const uint32_t* __restrict__ vals, const uint32_t* __restrict__ start,
                                                  uint32_t ncell, double* __restrict__ st) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    for (uint32_t c = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); c < ncell; c += waves) {
        const uint32_t s0 = start[c], s1 = start[c + 1];
        if (s0 == s1) continue;
        const float* lc = lobes + (size_t)c * A.K * 5;
        // every loop over the lobes is unrolled to kMaxLobes with a guard, so the sums stay
        // in registers (a loop bounded by K would index them dynamically, through scratch)
        double acc[kMaxLobes][4];
#pragma unroll
        for (int k = 0; k < kMaxLobes; ++k)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[k][j] = 0.0;
        for (uint32_t j = s0 + lane; j < s1; j += 64u) {
            f3 p, d;
            float w;
            src.get(vals[j], p, d, w);
            float g[kMaxLobes];
            float tot = 0.0f;
#pragma unroll
            for (int k = 0; k < kMaxLobes; ++k) {
                g[k] = 0.0f;
                if (k < (int)A.K) {
                    g[k] = lc[k * 5 + 4] * vmf_pdf(mk(lc[k * 5], lc[k * 5 + 1], lc[k * 5 + 2]), lc[k * 5 + 3], d);
                    tot += g[k];
                }
            }
#pragma unroll
            for (int k = 0; k < kMaxLobes; ++k) {
                if (k < (int)A.K) {
                    const float gk = (tot > 0.0f) ? g[k] / tot : 1.0f / (float)A.K;
                    const double wg = (double)w * (double)gk;
                    acc[k][0] += wg;
                    acc[k][1] += wg * (double)d.x;
                    acc[k][2] += wg * (double)d.y;
                    acc[k][3] += wg * (double)d.z;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < kMaxLobes; ++k) {
            if (k < (int)A.K) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    double v = acc[k][j];
#pragma unroll
                    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m);
                    acc[k][j] = v;
                }
            }
        }
        if (lane == 0) {
            double* sc = st + (size_t)c * A.K * 4;
#pragma unroll
            for (int k = 0; k < kMaxLobes; ++k) {
                if (k < (int)A.K) {
#pragma unroll
                    for (int j = 0; j < 4; ++j) sc[k * 4 + j] += acc[k][j];
                }
            }
        }
    }
}

// running stats = decay * running + frame; M-step; frame stats cleared
__global__ void k_vmm_mstep(uint32_t ncell, uint32_t K, float decay, float kappa_max, double min_weight,
                            double* __restrict__ run, double* __restrict__ frame, float* __restrict__ lobes,
                            uint32_t* __restrict__ valid) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    double* R = run + (size_t)c * K * 4;
    double* F = frame + (size_t)c * K * 4;
    double W = 0.0;
    for (uint32_t j = 0; j < K * 4; ++j) {
        R[j] = (double)decay * R[j] + F[j];
        F[j] = 0.0;
    }
    for (uint32_t k = 0; k < K; ++k) W += R[k * 4];
    if (!(W > min_weight)) return;
    float* lc = lobes + (size_t)c * K * 5;
    for (uint32_t k = 0; k < K; ++k) {
        double s0 = R[k * 4], sx = R[k * 4 + 1], sy = R[k * 4 + 2], sz = R[k * 4 + 3];
        double len = sqrt(sx * sx + sy * sy + sz * sz);
        if (!(s0 > 0.0)) { lc[k * 5 + 4] = 0.0f; continue; }
        double r = len / s0;
        r = r < 1e-4 ? 1e-4 : (r > 0.9999 ? 0.9999 : r);
        if (len > 1e-12) { lc[k * 5] = (float)(sx / len); lc[k * 5 + 1] = (float)(sy / len); lc[k * 5 + 2] = (float)(sz / len); }
        double kap = r * (3.0 - r * r) / (1.0 - r * r);
        kap = kap < 0.5 ? 0.5 : (kap > (double)kappa_max ? (double)kappa_max : kap);
        lc[k * 5 + 3] = (float)kap;
        lc[k * 5 + 4] = (float)(s0 / W);
    }
    valid[c] = 1u;
}

// initial lobes: cube-corner directions, kappa 2, equal weights (cells start invalid)
__global__ void k_vmm_init(uint32_t ncell, uint32_t K, float* __restrict__ lobes, uint32_t* __restrict__ valid) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    float* lc = lobes + (size_t)c * K * 5;
    const float s = 0.577350269f;
    for (uint32_t k = 0; k < K; ++k) {
        lc[k * 5 + 0] = (k & 1u) ? s : -s;
        lc[k * 5 + 1] = (k & 2u) ? s : -s;
        lc[k * 5 + 2] = (k & 4u) ? s : -s;
        lc[k * 5 + 3] = 2.0f;
        lc[k * 5 + 4] = 1.0f / (float)K;
    }
    valid[c] = 0u;
}

// lobes exposed to sampling: a cell is used only after its first M-step
__global__ void k_vmm_mask(uint32_t ncell, uint32_t K, const float* __restrict__ lobes, const uint32_t* __restrict__ valid,
                           float* __restrict__ out) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    for (uint32_t j = 0; j < K * 5; ++j) {
        float v = lobes[(size_t)c * K * 5 + j];
        if (!valid[c] && (j % 5u) == 4u) v = 0.0f;
        out[(size_t)c * K * 5 + j] = v;
    }
}

// ---- path-vertex finalisation: incident radiance along the sampled direction ----
// training sample: t0 (pos, lum(Li) / pdf)  t1 (wi, pdf)  t2 (n, lum(Li f cos) / pdf)  t3 (wo)
// The vertex's training sample, if it is kept (deterministic: a hash of the record index
// thins the live vertices to about the requested count).
__device__ __forceinline__ bool finalize_one(const float4* __restrict__ rec, uint32_t i, uint32_t max_bounces,
                                             uint32_t stamp, uint32_t shift, uint32_t phase,
                                             const float4* __restrict__ rad, uint32_t keep_thresh,
                                             float4& t0, float4& t1, float4& t2, float4& t3) {
    bool take = false;
    t0 = make_float4(0, 0, 0, 0); t1 = t0; t2 = t0; t3 = t0;
    const float4* r = rec + (size_t)i * kRecF4;
    float4 r0 = r[0];
    if (__float_as_uint(r0.w) == stamp && (pcg_hash(i ^ (stamp * 2654435761u)) >> 8) < keep_thresh) {
        float4 r1 = r[1], r2 = r[2], r3 = r[3], r4 = r[4], r5 = r[5], r6 = r[6], r7 = r[7];
        uint32_t pixel = ((i / max_bounces) << shift) + phase;
        float4 Lf = rad[pixel];
        const f3 nee = (r3.w > 0.5f) ? xyz(r5) : mk(0, 0, 0);
        f3 Ls = sub(sub(xyz(Lf), xyz(r4)), nee);
        // radiance gathered after this vertex, by subtraction: a remainder within the
        // rounding of the sums is zero (else dividing by a small throughput inflates it)
        const float tol = 1e-5f * (fabsf(Lf.x) + fabsf(Lf.y) + fabsf(Lf.z) + fabsf(r4.x) + fabsf(r4.y) +
                                   fabsf(r4.z) + fabsf(nee.x) + fabsf(nee.y) + fabsf(nee.z));
        Ls = mk(Ls.x > tol ? Ls.x : 0.0f, Ls.y > tol ? Ls.y : 0.0f, Ls.z > tol ? Ls.z : 0.0f);
        f3 tn = xyz(r6);
        f3 Li = mk(tn.x > 0.0f ? Ls.x / tn.x : 0.0f, tn.y > 0.0f ? Ls.y / tn.y : 0.0f, tn.z > 0.0f ? Ls.z / tn.z : 0.0f);
        float w = fmaxf(lum709(Li), 0.0f);
        const float wl = (r1.w > 0.0f) ? w / r1.w : 0.0f;
        if (wl > 0.0f && isfinite(wl)) {
            take = true;
            t0 = make_float4(r0.x, r0.y, r0.z, wl);
            t1 = make_float4(r3.x, r3.y, r3.z, r1.w);
            // product weight: Li f cos / pdf, the one-sample estimate of the integrand
            f3 bw = xyz(r7);
            const float wp = fmaxf(lum709(mk(Li.x * bw.x, Li.y * bw.y, Li.z * bw.z)), 0.0f);
            t2 = make_float4(r1.x, r1.y, r1.z, isfinite(wp) ? wp : 0.0f);
            t3 = make_float4(r2.x, r2.y, r2.z, 0.0f);
        }
    }
    return take;
}

// Compaction in record order, so the training set (and everything fitted from it) does not
// depend on scheduling: k_finalize_count counts the kept vertices of each 256-record chunk,
// an exclusive scan gives every chunk its first slot, k_finalize_write ranks the kept vertices
// inside the chunk (wave ballots + per-wave totals in LDS) and writes them; the last chunk
// stores the total.
constexpr uint32_t kFinChunk = 256;
__global__ void __launch_bounds__(kFinChunk) k_finalize_count(const float4* __restrict__ rec, uint32_t nrec,
                                                            uint32_t max_bounces, uint32_t stamp, uint32_t shift,
                                                            uint32_t phase, const float4* __restrict__ rad,
                                                            uint32_t keep_thresh, uint32_t* __restrict__ chunk_cnt) {
    const uint32_t nchunk = (nrec + kFinChunk - 1) / kFinChunk;
    for (uint32_t ch = blockIdx.x; ch < nchunk; ch += gridDim.x) {
        const uint32_t i = ch * kFinChunk + threadIdx.x;
        float4 t0, t1, t2, t3;
        const bool take = i < nrec && finalize_one(rec, i, max_bounces, stamp, shift, phase, rad, keep_thresh, t0, t1, t2, t3);
        const int n = __syncthreads_count(take ? 1 : 0);
        if (threadIdx.x == 0) chunk_cnt[ch] = (uint32_t)n;
    }
}

__global__ void __launch_bounds__(kFinChunk) k_finalize_write(const float4* __restrict__ rec, uint32_t nrec,
                                                            uint32_t max_bounces, uint32_t stamp, uint32_t shift,
                                                            uint32_t phase, const float4* __restrict__ rad,
                                                            uint32_t keep_thresh,
                                                            const uint32_t* __restrict__ chunk_cnt,
                                                            const uint32_t* __restrict__ chunk_off,
                                                            float4* __restrict__ train, uint32_t* __restrict__ cnt,
                                                            uint32_t cap) {
    __shared__ uint32_t s_wave[kFinChunk / 64];
    const uint32_t nchunk = (nrec + kFinChunk - 1) / kFinChunk;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    for (uint32_t ch = blockIdx.x; ch < nchunk; ch += gridDim.x) {
        const uint32_t i = ch * kFinChunk + threadIdx.x;
        float4 t0, t1, t2, t3;
        const bool take = i < nrec && finalize_one(rec, i, max_bounces, stamp, shift, phase, rad, keep_thresh, t0, t1, t2, t3);
        const unsigned long long m = __ballot(take ? 1 : 0);
        if (lane == 0) s_wave[wave] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t before = 0;
        for (uint32_t w = 0; w < wave; ++w) before += s_wave[w];
        const unsigned long long lt = (lane == 0u) ? 0ull : (~0ull >> (64u - lane));
        const uint32_t slot = chunk_off[ch] + before + (uint32_t)__popcll(m & lt);
        if (take && slot < cap) {
            train[(size_t)slot * 4] = t0; train[(size_t)slot * 4 + 1] = t1;
            train[(size_t)slot * 4 + 2] = t2; train[(size_t)slot * 4 + 3] = t3;
        }
        if (ch == nchunk - 1 && threadIdx.x == 0) *cnt = chunk_off[ch] + chunk_cnt[ch];
        __syncthreads();  // s_wave is rewritten by the next chunk
    }
}

VmmArgs vmm_args(const rrh_grid* g) {
    VmmArgs A;
    A.res = g->p.res; A.K = g->p.lobes;
    for (int k = 0; k < 3; ++k) { A.bmin[k] = g->p.bounds_min[k]; A.bext[k] = g->p.bounds_max[k] - g->p.bounds_min[k]; }
    return A;
}

// E-step over up to `cap` samples of src (the count may live on the device): cell keys,
// a stable radix sort of (key, index), the per-cell ranges, then the per-cell sums
template <class Src>
int vmm_estep_sorted(rrh_grid* g, Src src, uint32_t cap) {
    rrh_ctx* c = g->ctx;
    if (cap == 0) return RRH_OK;
    int rc;
    if ((rc = g->em_keys.ensure((size_t)cap * 8))) return rc;
    if ((rc = g->em_vals.ensure((size_t)cap * 8))) return rc;
    if ((rc = g->em_start.ensure((size_t)(g->ncell + 1) * 4))) return rc;
    uint32_t* k = (uint32_t*)g->em_keys.p;
    uint32_t* v = (uint32_t*)g->em_vals.p;
    // keys are cells in [0, ncell] (ncell = past the last cell): only their low bits are sorted
    const unsigned end_bit = 32u - (unsigned)__builtin_clz(g->ncell);
    size_t tb = 0;
    RRH_HIP(rocprim::radix_sort_pairs(nullptr, tb, k, k + cap, v, v + cap, cap, 0u, end_bit, c->stream));
    if ((rc = g->em_tmp.ensure(tb > 0 ? tb : 1))) return rc;
    const uint32_t gk = std::min((cap + 255u) / 256u, 8192u);
    hipLaunchKernelGGL(k_em_keys<Src>, dim3(gk), dim3(256), 0, c->stream, vmm_args(g), src, cap, g->ncell, k, v);
    RRH_LAUNCH_CHECK();
    tb = g->em_tmp.bytes;
    const hipError_t e = rocprim::radix_sort_pairs(g->em_tmp.p, tb, k, k + cap, v, v + cap, cap, 0u, end_bit, c->stream);
    if (e != hipSuccess) { set_error(std::string("EM sample sort: ") + hipGetErrorString(e)); return RRH_E_HIP; }
    const uint32_t gr = std::min((cap + 1u + 255u) / 256u, 8192u);
    hipLaunchKernelGGL(k_em_ranges, dim3(gr), dim3(256), 0, c->stream, (const uint32_t*)(k + cap), cap, g->ncell,
                       (uint32_t*)g->em_start.p);
    RRH_LAUNCH_CHECK();
    const uint32_t gc = std::min((g->ncell + 3u) / 4u, 4096u);
    hipLaunchKernelGGL(k_em_cells<Src>, dim3(gc), dim3(256), 0, c->stream, vmm_args(g), src, (const float*)g->lobes.p,
                       (const uint32_t*)(v + cap), (const uint32_t*)g->em_start.p, g->ncell, (double*)g->fstats.p);
    RRH_LAUNCH_CHECK();
    return RRH_OK;
}

int vmm_mstep(rrh_grid* g) {
    rrh_ctx* c = g->ctx;
    hipLaunchKernelGGL(k_vmm_mstep, dim3((g->ncell + 255) / 256), dim3(256), 0, c->stream, g->ncell, g->p.lobes,
                       g->p.decay, g->p.kappa_max, (double)g->p.min_weight, (double*)g->stats.p, (double*)g->fstats.p,
                       (float*)g->lobes.p, (uint32_t*)g->valid.p);
    RRH_LAUNCH_CHECK();
    return RRH_OK;
}

}  // namespace

namespace rrh {

int guide_prepare(rrh_ctx* c, rrh_scene* s, const rrh_pt_params* p, rrh_grid* grid, rrh_mlp* mlp, GuideDev& G,
                  VertexRec& R) {
    G = GuideDev();
    R = VertexRec();
    if (p->guide_mode == RRH_GUIDE_VMM) {
        RRH_ARG(grid && grid->p.kind == RRH_GRID_VMM, "guided render: RRH_GUIDE_VMM needs a VMM grid");
        int rc = c->scratch.ensure((size_t)grid->ncell * grid->p.lobes * 5 * 4);
        if (rc) return rc;
        hipLaunchKernelGGL(k_vmm_mask, dim3((grid->ncell + 255) / 256), dim3(256), 0, c->stream, grid->ncell,
                           grid->p.lobes, (const float*)grid->lobes.p, (const uint32_t*)grid->valid.p, (float*)c->scratch.p);
        RRH_LAUNCH_CHECK();
        G.lobes = (const float*)c->scratch.p;
        G.res = grid->p.res;
        G.K = grid->p.lobes;
        for (int k = 0; k < 3; ++k) { G.bmin[k] = grid->p.bounds_min[k]; G.bext[k] = grid->p.bounds_max[k] - grid->p.bounds_min[k]; }
    } else if (p->guide_mode == RRH_GUIDE_MLP) {
        RRH_ARG(mlp, "guided render: RRH_GUIDE_MLP needs an rrh_mlp");
        uint32_t npix = s->cam_host.screen_width * s->cam_host.screen_height;
        int rc = c->scratch.ensure((size_t)npix * kMlpLobeHalves * 2);
        if (rc) return rc;
        G.mlp_lobes = (const uint16_t*)c->scratch.p;
        G.K = kMaxLobes;
    }
    if (p->record_vertices) {
        uint32_t npix = s->cam_host.screen_width * s->cam_host.screen_height;
        size_t nrec = (size_t)npix * p->max_bounces;
        int rc = c->vtx.ensure(nrec * kRecF4 * 16);
        if (rc) return rc;
        c->stamp++;
        if (c->stamp == 0) c->stamp = 1;
        R.v = (float4*)c->vtx.p;
        R.max_bounces = p->max_bounces;
        R.stamp = c->stamp;
        // record a pixel lattice with at least four vertex slots per requested vertex (paths end
        // early, so only part of the slots hold vertices), so the hash thinning still picks
        // the training set
        const double want = 4.0 * (double)p->record_vertices;
        while (R.shift < 16 && (double)(npix >> (R.shift + 1)) * p->max_bounces >= want) ++R.shift;
        R.phase = c->stamp & ((1u << R.shift) - 1u);
        c->last_rec_bounces = p->max_bounces;
    }
    return RRH_OK;
}

int guide_finish(rrh_ctx* c, rrh_scene* s, const rrh_pt_params* p, rrh_grid* grid, rrh_mlp* mlp, const VertexRec& R,
                 const float4* rad) {
    (void)grid; (void)mlp;
    if (!R.v) return RRH_OK;
    uint32_t npix = s->cam_host.screen_width * s->cam_host.screen_height;
    uint32_t nrec = ((npix + (1u << R.shift) - 1u) >> R.shift) * p->max_bounces;  // the recorded lattice
    uint32_t target = p->record_vertices;
    uint32_t cap = target;
    int rc = c->train.ensure((size_t)cap * 64);
    if (rc) return rc;
    if ((rc = c->train_cnt.ensure(4))) return rc;
    c->train_capacity = cap;
    // keep about `target` of the (roughly nrec/2) live vertices
    double keep = (double)target / ((double)nrec * 0.5);
    uint32_t thresh = keep >= 1.0 ? (1u << 24) : (uint32_t)(keep * (double)(1u << 24));
    if (nrec == 0) {
        RRH_HIP(hipMemsetAsync(c->train_cnt.p, 0, 4, c->stream));
        return RRH_OK;
    }
    const uint32_t nchunk = (nrec + kFinChunk - 1) / kFinChunk;
    if ((rc = c->fin_chunks.ensure((size_t)nchunk * 8))) return rc;
    uint32_t* ccnt = (uint32_t*)c->fin_chunks.p;
    uint32_t* coff = ccnt + nchunk;
    size_t tb = 0;
    RRH_HIP(rocprim::exclusive_scan(nullptr, tb, ccnt, coff, 0u, nchunk, rocprim::plus<uint32_t>(), c->stream));
    if ((rc = c->fin_tmp.ensure(tb > 0 ? tb : 1))) return rc;
    const uint32_t g = std::min(nchunk, 8192u);
    hipLaunchKernelGGL(k_finalize_count, dim3(g), dim3(kFinChunk), 0, c->stream, (const float4*)c->vtx.p, nrec,
                       p->max_bounces, R.stamp, R.shift, R.phase, rad, thresh, ccnt);
    RRH_LAUNCH_CHECK();
    tb = c->fin_tmp.bytes;
    const hipError_t e = rocprim::exclusive_scan(c->fin_tmp.p, tb, ccnt, coff, 0u, nchunk, rocprim::plus<uint32_t>(),
                                                 c->stream);
    if (e != hipSuccess) { set_error(std::string("vertex compaction scan: ") + hipGetErrorString(e)); return RRH_E_HIP; }
    hipLaunchKernelGGL(k_finalize_write, dim3(g), dim3(kFinChunk), 0, c->stream, (const float4*)c->vtx.p, nrec,
                       p->max_bounces, R.stamp, R.shift, R.phase, rad, thresh, (const uint32_t*)ccnt, (const uint32_t*)coff,
                       (float4*)c->train.p, (uint32_t*)c->train_cnt.p, cap);
    RRH_LAUNCH_CHECK();
    return RRH_OK;
}

}  // namespace rrh

extern "C" {

int rrh_grid_create(rrh_ctx* c, const rrh_grid_params* p, rrh_grid** out) {
    RRH_ARG(c && p && out, "rrh_grid_create: null argument");
    *out = nullptr;
    rrh_grid* g = new rrh_grid();
    g->ctx = c;
    g->p = *p;
    int rc = RRH_OK;
    if (p->kind == RRH_GRID_HASH) {
        if (p->cell_count == 0 || p->cell_count > (1u << 26)) { delete g; set_error("rrh_grid_create: cell_count must be in [1, 2^26]"); return RRH_E_ARG; }
        g->ncell = p->cell_count;
        if (!rc) rc = g->accum.ensure((size_t)g->ncell * 16);
        if (!rc) rc = g->fitted.ensure((size_t)g->ncell * 16);
    } else if (p->kind == RRH_GRID_VMM) {
        if (p->res == 0 || p->res > 128 || p->lobes == 0 || p->lobes > kMaxLobes) { delete g; set_error("rrh_grid_create: res in [1,128], lobes in [1,8]"); return RRH_E_ARG; }
        for (int k = 0; k < 3; ++k)
            if (!(p->bounds_max[k] > p->bounds_min[k])) { delete g; set_error("rrh_grid_create: empty bounds"); return RRH_E_ARG; }
        g->ncell = p->res * p->res * p->res;
        size_t nl = (size_t)g->ncell * p->lobes;
        if (!rc) rc = g->lobes.ensure(nl * 5 * 4);
        if (!rc) rc = g->stats.ensure(nl * 4 * 8);
        if (!rc) rc = g->fstats.ensure(nl * 4 * 8);
        if (!rc) rc = g->valid.ensure((size_t)g->ncell * 4);
    } else {
        delete g;
        set_error("rrh_grid_create: unknown kind");
        return RRH_E_ARG;
    }
    if (rc) { rrh_grid_destroy(g); return rc; }
    rc = rrh_grid_reset(c, g);
    if (rc) { rrh_grid_destroy(g); return rc; }
    *out = g;
    return RRH_OK;
}

int rrh_grid_destroy(rrh_grid* g) {
    if (!g) return RRH_OK;
    (void)hipDeviceSynchronize();  // the owning context may already be gone
    g->accum.release(); g->fitted.release(); g->lobes.release(); g->stats.release(); g->fstats.release(); g->valid.release();
    g->em_keys.release(); g->em_vals.release(); g->em_tmp.release(); g->em_start.release();
    delete g;
    return RRH_OK;
}

// ReSTIRPTPass.cpp:602-612 clears both buffers when the grid is dirty
int rrh_grid_reset(rrh_ctx* c, rrh_grid* g) {
    RRH_ARG(c && g, "rrh_grid_reset: null argument");
    if (g->p.kind == RRH_GRID_HASH) {
        RRH_HIP(hipMemsetAsync(g->accum.p, 0, g->accum.bytes, c->stream));
        RRH_HIP(hipMemsetAsync(g->fitted.p, 0, g->fitted.bytes, c->stream));
    } else {
        RRH_HIP(hipMemsetAsync(g->stats.p, 0, g->stats.bytes, c->stream));
        RRH_HIP(hipMemsetAsync(g->fstats.p, 0, g->fstats.bytes, c->stream));
        hipLaunchKernelGGL(k_vmm_init, dim3((g->ncell + 255) / 256), dim3(256), 0, c->stream, g->ncell, g->p.lobes,
                           (float*)g->lobes.p, (uint32_t*)g->valid.p);
        RRH_LAUNCH_CHECK();
    }
    return RRH_OK;
}

static int upload_tmp(rrh_ctx* c, const void* src, size_t bytes, void** dst) {
    RRH_HIP(hipMalloc(dst, bytes ? bytes : 16));
    if (bytes) RRH_HIP(hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    return RRH_OK;
}

int rrh_grid_splat(rrh_ctx* c, rrh_grid* g, const float* pos, const uint32_t* bsdf, const float* dirs, uint32_t n,
                   const float cam_pos[3]) {
    RRH_ARG(c && g && pos && bsdf && dirs && cam_pos, "rrh_grid_splat: null argument");
    RRH_ARG(g->p.kind == RRH_GRID_HASH, "rrh_grid_splat: HASH grids only");
    if (n == 0) return RRH_OK;
    void *dp = nullptr, *db = nullptr, *dd = nullptr;
    int rc = upload_tmp(c, pos, (size_t)n * 12, &dp);
    if (!rc) rc = upload_tmp(c, bsdf, (size_t)n * 4, &db);
    if (!rc) rc = upload_tmp(c, dirs, (size_t)n * 12, &dd);
    if (!rc) {
        hipLaunchKernelGGL(k_hash_splat, dim3((n + 255) / 256), dim3(256), 0, c->stream, (const float*)dp,
                           (const uint32_t*)db, (const float*)dd, n, mk(cam_pos[0], cam_pos[1], cam_pos[2]),
                           g->p.base_cell_size, g->p.distance_slope, g->p.cell_count, (uint32_t*)g->accum.p);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) { set_error(hipGetErrorString(e)); rc = RRH_E_HIP; }
    }
    (void)hipFree(dp); (void)hipFree(db); (void)hipFree(dd);
    return rc;
}

int rrh_grid_fit(rrh_ctx* c, rrh_grid* g) {
    RRH_ARG(c && g, "rrh_grid_fit: null argument");
    if (g->p.kind == RRH_GRID_HASH) {
        hipLaunchKernelGGL(k_hash_fit, dim3((g->ncell + 63) / 64), dim3(64), 0, c->stream, (uint4*)g->accum.p,
                           (float4*)g->fitted.p, g->ncell, g->p.min_samples, g->p.kappa_max);
        RRH_LAUNCH_CHECK();
        return RRH_OK;
    }
    return vmm_mstep(g);
}

int rrh_grid_sample_primary(rrh_ctx* c, rrh_grid* g, const float* pos, const float* nrm, const uint32_t* bsdf,
                            uint32_t* seeds, uint32_t n, const float cam_pos[3], float alpha, float* dirs, float* pdfs) {
    RRH_ARG(c && g && pos && nrm && bsdf && seeds && cam_pos && dirs && pdfs, "rrh_grid_sample_primary: null argument");
    RRH_ARG(g->p.kind == RRH_GRID_HASH, "rrh_grid_sample_primary: HASH grids only");
    if (n == 0) return RRH_OK;
    void *dp = nullptr, *dn = nullptr, *db = nullptr, *ds = nullptr, *dd = nullptr, *dq = nullptr;
    int rc = upload_tmp(c, pos, (size_t)n * 12, &dp);
    if (!rc) rc = upload_tmp(c, nrm, (size_t)n * 12, &dn);
    if (!rc) rc = upload_tmp(c, bsdf, (size_t)n * 4, &db);
    if (!rc) rc = upload_tmp(c, seeds, (size_t)n * 4, &ds);
    if (!rc) rc = upload_tmp(c, nullptr, 0, &dd);
    if (!rc) { (void)hipFree(dd); rc = (hipMalloc(&dd, (size_t)n * 12) == hipSuccess) ? RRH_OK : RRH_E_NOMEM; }
    if (!rc) rc = (hipMalloc(&dq, (size_t)n * 4) == hipSuccess) ? RRH_OK : RRH_E_NOMEM;
    if (!rc) {
        hipLaunchKernelGGL(k_hash_sample, dim3((n + 255) / 256), dim3(256), 0, c->stream, (const float4*)g->fitted.p,
                           (const float*)dp, (const float*)dn, (const uint32_t*)db, (uint32_t*)ds, n,
                           mk(cam_pos[0], cam_pos[1], cam_pos[2]), g->p.base_cell_size, g->p.distance_slope,
                           g->p.cell_count, alpha, (float*)dd, (float*)dq);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(dirs, dd, (size_t)n * 12, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(pdfs, dq, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(seeds, ds, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) { set_error(hipGetErrorString(e)); rc = RRH_E_HIP; }
    }
    (void)hipFree(dp); (void)hipFree(dn); (void)hipFree(db); (void)hipFree(ds); (void)hipFree(dd); (void)hipFree(dq);
    return rc;
}

int rrh_grid_em_step(rrh_ctx* c, rrh_grid* g, const float* pos, const float* dirs, const float* w, uint32_t n) {
    RRH_ARG(c && g && pos && dirs && w, "rrh_grid_em_step: null argument");
    RRH_ARG(g->p.kind == RRH_GRID_VMM, "rrh_grid_em_step: VMM grids only");
    void *dp = nullptr, *dd = nullptr, *dw = nullptr;
    int rc = upload_tmp(c, pos, (size_t)n * 12, &dp);
    if (!rc) rc = upload_tmp(c, dirs, (size_t)n * 12, &dd);
    if (!rc) rc = upload_tmp(c, w, (size_t)n * 4, &dw);
    if (!rc) rc = vmm_estep_sorted(g, ArraySrc{(const float*)dp, (const float*)dd, (const float*)dw}, n);
    if (!rc) rc = vmm_mstep(g);
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) { set_error("rrh_grid_em_step: sync failed"); rc = RRH_E_HIP; }
    (void)hipFree(dp); (void)hipFree(dd); (void)hipFree(dw);
    return rc;
}

int rrh_grid_em_from_vertices(rrh_ctx* c, rrh_grid* g) {
    RRH_ARG(c && g, "rrh_grid_em_from_vertices: null argument");
    RRH_ARG(g->p.kind == RRH_GRID_VMM, "rrh_grid_em_from_vertices: VMM grids only");
    RRH_ARG(c->train.p && c->train_cnt.p, "rrh_grid_em_from_vertices: no recorded vertices (render with record_vertices)");
    const int rc = vmm_estep_sorted(g, TrainSrc{(const float4*)c->train.p, (const uint32_t*)c->train_cnt.p},
                                    c->train_capacity);
    return rc ? rc : vmm_mstep(g);
}

int rrh_grid_download(rrh_ctx* c, const rrh_grid* g, void* a, size_t b0, void* b, size_t b1) {
    RRH_ARG(c && g, "rrh_grid_download: null argument");
    const DevBuf& A = (g->p.kind == RRH_GRID_HASH) ? g->fitted : g->lobes;
    const DevBuf& B = (g->p.kind == RRH_GRID_HASH) ? g->accum : g->stats;
    if (a) { RRH_ARG(b0 == A.bytes, "rrh_grid_download: first buffer size mismatch"); RRH_HIP(hipMemcpyAsync(a, A.p, b0, hipMemcpyDeviceToHost, c->stream)); }
    if (b) { RRH_ARG(b1 == B.bytes, "rrh_grid_download: second buffer size mismatch"); RRH_HIP(hipMemcpyAsync(b, B.p, b1, hipMemcpyDeviceToHost, c->stream)); }
    RRH_HIP(hipStreamSynchronize(c->stream));
    return RRH_OK;
}

int rrh_grid_upload(rrh_ctx* c, rrh_grid* g, const void* a, size_t b0) {
    RRH_ARG(c && g && a, "rrh_grid_upload: null argument");
    DevBuf& A = (g->p.kind == RRH_GRID_HASH) ? g->fitted : g->lobes;
    RRH_ARG(b0 == A.bytes, "rrh_grid_upload: size mismatch");
    RRH_HIP(hipMemcpyAsync(A.p, a, b0, hipMemcpyHostToDevice, c->stream));
    if (g->p.kind == RRH_GRID_VMM) {
        std::vector<uint32_t> ones(g->ncell, 1u);
        RRH_HIP(hipMemcpyAsync(g->valid.p, ones.data(), (size_t)g->ncell * 4, hipMemcpyHostToDevice, c->stream));
    }
    RRH_HIP(hipStreamSynchronize(c->stream));
    return RRH_OK;
}

}  // extern "C"

// rrh_pt.hip -- wavefront path tracer (replaces pathtracer.slang cs_main :181-265 and
// PathTracerPass::execute, src/passes/pathtracer_reference/PathTracerPass.cpp:170-224)
// plus progressive accumulation (accumulate.slang:21-46, AccumulatePass.cpp:167-218).
//
// One frame = raygen, then per bounce: closest-hit traversal over the live ray
// queue, shading (NEE candidate + BSDF sample + Russian roulette) that appends
// surviving paths to the next queue and shadow rays to the shadow queue, and an
// occlusion pass that adds unshadowed NEE.  Each path keeps the reference's
// per-pixel RNG stream, so the draw order per pixel is the reference's:
// jitter(2) | per bounce: [light pick (+2 for a triangle light)] [BSDF 2] [RR 1].
// Queues are compacted with one wave-aggregated atomic per wave.
#include <cstring>
#include <mutex>
#include <rocprim/device/device_radix_sort.hpp>

#include "rrh_internal.h"
#include "rrh_guide_dev.h"

using namespace rrh;

namespace {

constexpr int kMaxBounces = 32;
constexpr uint32_t kBlock = 256;
// the ray-queue traversal kernels (k_intersect, k_shadow): one-wave workgroups, one ray per
// lane (measured on the 1080p frame: 294 us per k_intersect launch against 311 us with
// 256-lane groups tracing two rays per lane; a workgroup's LDS is held until its slowest wave
// ends, so smaller groups free it sooner)
constexpr uint32_t kTraceBlock = 64;
constexpr uint32_t kTraceGridCap = 32768;

// counters layout (uint32): [0, 64) ray-queue size per bounce, [64, 128) shadow rays per bounce,
// [128, 192) / [192, 256) next queue slot handed out by the persistent waves of bounce b's
// closest-hit launch (with its paired shadow queue) / of a launch of bounce b's shadow rays alone,
// or, in the suspend mode, the rays those launches suspended
struct Q {
    float4 *o, *d, *thp;
};

// the frame's device scratch: ray queues (ping-pong), hits, shadow queue, counters
struct Bufs {
    Q q[2];
    float4 *hits, *sh_o, *sh_d, *sh_c;
    uint32_t* cnt;
};

// Scratch of rrh_pt_render_noalloc: the same ten float4 planes and the counters, for frames of
// up to kNoAllocMaxPixels, in a file-scope device array (each GPU holds its own copy until the
// process ends).  Host code uses the per-device address from hipGetSymbolAddress, never the
// array's name.
constexpr uint32_t kNoAllocMaxPixels = 512u * 512u;
constexpr size_t kNoAllocPlane = (size_t)kNoAllocMaxPixels * 16u;
constexpr size_t kNoAllocBytes = 10u * kNoAllocPlane + 256u * 4u;
__device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_addr[kMaxDevices] = {};

__global__ void k_zero_u32(uint32_t* __restrict__ p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// one 8x8 pixel tile per wave; the queue is filled in tile order (no atomics): for an
// image whose sides are multiples of 8 the slot is tile * 64 + lane, otherwise row-major
// the primary rays of image rows [r0, r0 + rows), 8x8 tiles in queue order
__global__ void __launch_bounds__(64) k_raygen(DevScene S, Q out, uint32_t* __restrict__ cnt, float4* __restrict__ rad,
                                               uint32_t frame_seed, uint32_t r0, uint32_t rows) {
    const rrh_camera& cam = *S.cam;
    uint32_t W = cam.screen_width;
    uint32_t px = blockIdx.x * 8 + (threadIdx.x & 7u);
    uint32_t ly = blockIdx.y * 8 + (threadIdx.x >> 3);  // row within the band
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) cnt[0] = W * rows;
    if (px >= W || ly >= rows) return;
    const uint32_t py = r0 + ly;
    uint32_t seed = make_seed(px, py, frame_seed);
    f3 d = primary_dir(cam, px, py, seed);
    uint32_t pix = py * W + px;
    rad[pix] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    uint32_t slot = ((W & 7u) == 0u && (rows & 7u) == 0u) ? (blockIdx.y * gridDim.x + blockIdx.x) * 64u + threadIdx.x
                                                         : ly * W + px;
    out.o[slot] = make_float4(cam.position[0], cam.position[1], cam.position[2], __uint_as_float(pix));
    out.d[slot] = make_float4(d.x, d.y, d.z, __uint_as_float(seed));
    out.thp[slot] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
}

// block-aggregated queue append: one atomic per block; every thread of the block calls it
__device__ __forceinline__ uint32_t block_push(uint32_t* counter, bool pred, uint32_t* s_cnt) {
    const uint32_t wid = threadIdx.x >> 6, lane = threadIdx.x & 63u, nw = blockDim.x >> 6;
    unsigned long long mask = __ballot(pred ? 1 : 0);
    unsigned long long lt = (lane == 0u) ? 0ull : (~0ull >> (64u - lane));
    uint32_t off = (uint32_t)__popcll(mask & lt);
    if (lane == 0) s_cnt[wid] = (uint32_t)__popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (uint32_t w = 0; w < nw; ++w) { uint32_t c = s_cnt[w]; s_cnt[w] = tot; tot += c; }
        s_cnt[nw] = tot ? atomicAdd(counter, tot) : 0u;
    }
    __syncthreads();
    uint32_t slot = s_cnt[nw] + s_cnt[wid] + off;
    __syncthreads();  // s_cnt is reused by the next call
    return slot;
}

// the top of the tree staged in LDS (TOP): dynamic shared memory of top_n x 64 B
extern __shared__ uint4 s_top[];

// closest hits of the extension queue's first n slots, workgroup blk of nblk (one wave each)
// perm (optional): the queue slots in tracing order (the ray sort); results stay in slot order
// PREP: the learned guide's input record of each slot as well (write_hit_prep into prep), in
// place of a separate input pass over the queue; a variant of its own, so the kernels without it
// keep their register count
template <bool COUNT, bool PREP, class Stk, class Top>
__device__ __forceinline__ void trace_ext_part(const DevScene& S, const Q& in, uint32_t n, uint32_t blk, uint32_t nblk,
                                               float4* __restrict__ hits, const uint32_t* __restrict__ perm,
                                               float4* __restrict__ prep, Stk& st, const Top& top, uint32_t& nn,
                                               uint32_t& nt, uint32_t& nr) {
    for (uint32_t j = blk * kTraceBlock + threadIdx.x; j < n; j += nblk * kTraceBlock) {
        const uint32_t i = perm ? perm[j] : j;
        if (i >= n) continue;  // a sorted order holds only live slots; never trust it past the queue
        float4 o = in.o[i], d = in.d[i];
        Hit h = trace_closest_s<COUNT>(S, mk(o.x, o.y, o.z), mk(d.x, d.y, d.z), 0.001f, 1e9f, st, &nn, &nt, top);
        const float4 h4 = make_float4(h.t, __uint_as_float(h.prim), h.b1, h.b2);
        hits[i] = h4;
        if constexpr (PREP) write_hit_prep(S, d, h4, prep, i);
        nr++;
    }
}
// shadow rays of bounce b (the queue's first n): a visible one adds its contribution to its pixel
template <bool COUNT, class Stk, class Top>
__device__ __forceinline__ void trace_shadow_part(const DevScene& S, uint32_t n, uint32_t b, uint32_t blk, uint32_t nblk,
                                                  const float4* __restrict__ sh_o, const float4* __restrict__ sh_d,
                                                  const float4* __restrict__ sh_c, float4* __restrict__ rad,
                                                  const VertexRec& rec, Stk& st, const Top& top, uint32_t& nn,
                                                  uint32_t& nt, uint32_t& nr) {
    for (uint32_t i = blk * kTraceBlock + threadIdx.x; i < n; i += nblk * kTraceBlock) {
        float4 o = sh_o[i], d = sh_d[i];
        nr++;
        if (!trace_occluded_s<COUNT>(S, mk(o.x, o.y, o.z), mk(d.x, d.y, d.z), 0.001f, o.w, st, top, &nn, &nt)) {
            uint32_t pixel = __float_as_uint(d.w);
            float4 c = sh_c[i];
            float4 L = rad[pixel];
            L.x = L.x + c.x; L.y = L.y + c.y; L.z = L.z + c.z;
            rad[pixel] = L;
            if (rec.v) guide_record_nee_visible(rec, pixel, b);
        }
    }
}

// per wave of the instrumented kernels: 64 x the largest per-lane node and triangle counts into
// w[0], w[1] (one ray per lane at the grid sizes used), so sum / (64 x max) over the waves bounds
// the lane utilisation that the spread of ray lengths alone allows
__device__ __forceinline__ void wave_count_max(uint32_t nn, uint32_t nt, unsigned long long* w) {
    for (int off = 32; off > 0; off >>= 1) {
        nn = max(nn, (uint32_t)__shfl_xor((int)nn, off));
        nt = max(nt, (uint32_t)__shfl_xor((int)nt, off));
    }
    if ((threadIdx.x & 63u) == 0u) {
        atomicAdd(w + 0, 64ull * nn);
        atomicAdd(w + 1, 64ull * nt);
    }
}

// TOP: the first top_n quantized nodes are read from LDS; PREP: see trace_ext_part
template <bool COUNT, int STK, bool TOP, bool PREP = false>
__global__ void __launch_bounds__(kTraceBlock) k_intersect(DevScene S, Q in, const uint32_t* __restrict__ cnt, uint32_t b,
                                                           float4* __restrict__ hits, unsigned long long* __restrict__ fetch,
                                                           bool remap, const uint32_t* __restrict__ perm, uint32_t top_n,
                                                           float4* __restrict__ prep) {
    __shared__ int s_stack[(kTraceBlock / 64) * StackSel<STK>::kLds * 64];
    const auto top = stage_top<TOP>(S, s_top, top_n);
    uint32_t nn = 0, nt = 0, nr = 0;
    auto st = make_stack<STK>(s_stack);
    trace_ext_part<COUNT, PREP>(S, in, cnt[b], xcd_block(blockIdx.x, gridDim.x, remap), gridDim.x, hits, perm, prep, st, top,
                                nn, nt, nr);
    if constexpr (COUNT) {
        atomicAdd(fetch + 0, (unsigned long long)nn);
        atomicAdd(fetch + 1, (unsigned long long)nt);
        atomicAdd(fetch + 2, (unsigned long long)nr);
        wave_count_max(nn, nt, fetch + 6);
    }
}

// Persistent variants (binary nodes): every wave loops, handing its finished lanes new rays
// from the queue (one atomic per refill, when at least kRefill lanes are idle), so lanes
// whose rays end early do not idle until the slowest ray of their wave is done.
constexpr uint32_t kRefill = 16;

template <bool COUNT, int STK>
__global__ void __launch_bounds__(kBlock) k_intersect_p(DevScene S, Q in, const uint32_t* __restrict__ cnt, uint32_t b,
                                                        float4* __restrict__ hits, uint32_t* __restrict__ next,
                                                        unsigned long long* __restrict__ fetch) {
    __shared__ int s_stack[(kBlock / 64) * StackSel<STK>::kLds * 64];
    auto st = make_stack<STK>(s_stack);
    const uint32_t n = cnt[b];
    uint32_t nn = 0, nt = 0, nr = 0;
    TravState T;
    uint32_t idx = 0;
    bool active = false, drained = false;
    while (true) {
        const bool idle = !active;
        if (!drained && __popcll(__ballot(idle ? 1 : 0)) >= (int)kRefill) {
            uint32_t my = 0;
            if (wave_refill(next, n, idle, drained, my)) {
                const float4 o = in.o[my], d = in.d[my];
                trav_begin(T, mk(o.x, o.y, o.z), mk(d.x, d.y, d.z), 0.001f, 1e9f);
                idx = my;
                active = true;
                nr++;
            }
        }
        if (__ballot(active ? 1 : 0) == 0ull) {
            if (drained) break;
            continue;
        }
        if (active && closest_step<COUNT>(S, T, st, nn, nt)) {
            hits[idx] = make_float4(T.t, __uint_as_float(T.prim), T.b1, T.b2);
            active = false;
        }
    }
    if constexpr (COUNT) {
        atomicAdd(fetch + 0, (unsigned long long)nn);
        atomicAdd(fetch + 1, (unsigned long long)nt);
        atomicAdd(fetch + 2, (unsigned long long)nr);
    }
}

template <int STK>
__global__ void __launch_bounds__(kBlock) k_shadow_p(DevScene S, const uint32_t* __restrict__ cnt, uint32_t b,
                                                     const float4* __restrict__ sh_o, const float4* __restrict__ sh_d,
                                                     const float4* __restrict__ sh_c, float4* __restrict__ rad,
                                                     VertexRec rec, uint32_t* __restrict__ next) {
    __shared__ int s_stack[(kBlock / 64) * StackSel<STK>::kLds * 64];
    auto st = make_stack<STK>(s_stack);
    const uint32_t n = cnt[64 + b];
    TravState T;
    uint32_t idx = 0;
    bool active = false, drained = false;
    while (true) {
        const bool idle = !active;
        if (!drained && __popcll(__ballot(idle ? 1 : 0)) >= (int)kRefill) {
            uint32_t my = 0;
            if (wave_refill(next, n, idle, drained, my)) {
                const float4 o = sh_o[my], d = sh_d[my];
                trav_begin(T, mk(o.x, o.y, o.z), mk(d.x, d.y, d.z), 0.001f, o.w);
                if (!(o.w >= 0.001f)) T.prim = 1u;  // empty extent: never occluded
                idx = my;
                active = true;
            }
        }
        if (__ballot(active ? 1 : 0) == 0ull) {
            if (drained) break;
            continue;
        }
        if (active) {
            bool done;
            if (T.prim == 1u) done = true;  // the empty-extent ray
            else done = occluded_step(S, T, st);
            if (done) {
                if (T.prim != 0u) {  // visible: add the NEE contribution
                    const float4 d = sh_d[idx], c = sh_c[idx];
                    const uint32_t pixel = __float_as_uint(d.w);
                    float4 L = rad[pixel];
                    L.x = L.x + c.x; L.y = L.y + c.y; L.z = L.z + c.z;
                    rad[pixel] = L;
                    if (rec.v) guide_record_nee_visible(rec, pixel, b);
                }
                active = false;
            }
        }
    }
}

// adds this frame's per-bounce queue sizes into the context's running totals
__global__ void k_sum_counts(const uint32_t* __restrict__ cnt, uint32_t max_bounces,
                             unsigned long long* __restrict__ tot) {
    if (threadIdx.x != 0) return;
    unsigned long long e = 0, sh = 0;
    for (uint32_t b = 1; b < max_bounces; ++b) e += cnt[b];
    for (uint32_t b = 0; b < max_bounces; ++b) sh += cnt[64 + b];
    tot[0] += cnt[0];
    tot[1] += e;
    tot[2] += sh;
    tot[3] += cnt[0] + e;  // the path tracer's own closest-hit and shadow rays (the DI pass adds to [0] and [2])
    tot[4] += sh;
}

struct NeeOut { bool shadow; f3 c; f3 dir; float tmax; };

// pathtracer.slang:124-177 up to (excluding) the shadow test; triangle lights
// use area sampling and the power heuristic against the BSDF sampling pdf.
__device__ __forceinline__ NeeOut nee_candidate(const DevScene& S, f3 hit, f3 n, f3 wo, const Mat& m, uint32_t& seed) {
    NeeOut r; r.shadow = false; r.c = mk(0, 0, 0); r.dir = mk(0, 0, 0); r.tmax = 0.0f;
    uint32_t N = S.nl + S.ntl;
    if (N == 0) return r;
    uint32_t li = pick_index(rnd(seed), N);
    if (li < S.nl) {
        const rrh_light& L = S.lights[li];
        f3 ldir; float dist, pl;
        if (L.type == 0) {
            f3 to = sub(ld3(L.position), hit);
            dist = length(to);
            ldir = divs(to, dist);
            pl = dist * dist;
        } else {
            ldir = normalize(neg(ld3(L.direction)));
            dist = 1e9f;
            pl = 1.0f;
        }
        if (dot(ldir, n) <= 0.0f) return r;
        f3 f = mat_eval(m, n, wo, ldir);
        float ndl = fmaxf(dot(n, ldir), 0.0f);
        f3 E = scale(ld3(L.emission), L.intensity);
        float den = (L.type == 0) ? (pl + 1e-6f) : 1.0f;
        f3 irr = divs(E, den);
        float pb = mat_ref_pdf(m, wo, ldir, n);
        float mis = (L.type == 0) ? 1.0f : (pl * pl) / (pl * pl + pb * pb + 1e-10f);
        r.c = scale(scale(scale(mul(f, irr), ndl), mis), (float)N);
        r.dir = ldir;
        r.tmax = dist - 0.01f;
        r.shadow = true;
        return r;
    }
    float u1 = rnd(seed);
    float u2 = rnd(seed);
    TriSample ts = tri_light_point(S, li - S.nl, u1, u2);
    f3 to = sub(ts.p, hit);
    float d2 = dot(to, to);
    float dist = sqrtf(d2);
    f3 l = scale(to, 1.0f / dist);
    if (dot(l, n) <= 0.0f) return r;
    float cosl = fabsf(dot(ts.nl, l));
    if (!(cosl > 0.0f)) return r;
    f3 f = mat_eval(m, n, wo, l);
    float ndl = fmaxf(dot(n, l), 0.0f);
    float pl = d2 / (cosl * ts.area) / (float)N;
    float pb = mat_sample_pdf(m, wo, l, n);
    float mis = (pl * pl) / (pl * pl + pb * pb);
    r.c = scale(scale(mul(f, ts.Le), ndl), mis / pl);
    r.dir = l;
    r.tmax = dist - 0.01f;
    r.shadow = true;
    return r;
}

struct ShadeArgs {
    uint32_t b, max_bounces, rr_start, guide_mode;
    bool remap;
    float guide_alpha;
    GuideDev guide;
    VertexRec rec;
    // when set: per queue slot, the hit's (pos, 1), (normal facing the ray origin, material index),
    // (wo, 0) as k_mlp_prep wrote them from fetch_hit (the same floats), in place of the lookup
    const float4* hit_prep = nullptr;
};

__global__ void __launch_bounds__(kBlock) k_shade(DevScene S, Q in, Q out, const float4* __restrict__ hits,
                                                  uint32_t* __restrict__ cnt, float4* __restrict__ sh_o,
                                                  float4* __restrict__ sh_d, float4* __restrict__ sh_c,
                                                  float4* __restrict__ rad, ShadeArgs A) {
    const uint32_t b = A.b;
    const uint32_t n = cnt[b];
    const uint32_t stride = gridDim.x * blockDim.x;
    const bool last = (b + 1u >= A.max_bounces);
    __shared__ uint32_t s_cnt[kBlock / 64 + 1];
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x, A.remap);
    for (uint32_t base = blk * blockDim.x; base < n; base += stride) {
        uint32_t i = base + threadIdx.x;
        bool valid = i < n;
        bool push_ext = false, push_sh = false;
        float4 no = make_float4(0, 0, 0, 0), nd = no, nt = no, so = no, sd = no, sc = no;
        if (valid) {
            float4 o4 = in.o[i], d4 = in.d[i], t4 = in.thp[i], h4 = hits[i];
            uint32_t pixel = __float_as_uint(o4.w);
            uint32_t seed = __float_as_uint(d4.w);
            f3 thp = xyz(t4);
            float last_pdf = t4.w;
            f3 dir = xyz(d4);
            uint32_t prim = __float_as_uint(h4.y);
            float4 L4 = rad[pixel];
            f3 L = xyz(L4);
            if (prim == ~0u) {
                L = add(L, mul(thp, sky(dir)));
                rad[pixel] = make_float4(L.x, L.y, L.z, L4.w);
            } else {
                HitInfo hi;
                if (A.hit_prep) {
                    const float4 p0 = A.hit_prep[(size_t)i * 3], p1 = A.hit_prep[(size_t)i * 3 + 1];
                    hi.pos = xyz(p0);
                    hi.normal = xyz(p1);
                    hi.mat = __float_as_uint(p1.w);
                } else {
                    hi = fetch_hit(S, prim, h4.z, h4.w);
                }
                Mat m = load_mat(S, hi.mat);
                if (is_emissive(m)) {
                    uint32_t lo = S.tri_light[prim];
                    if (b > 0 && lo != ~0u) {
                        float pl = tri_light_pdf(S, lo - S.nl, dir, h4.x);
                        float w = (last_pdf * last_pdf) / (last_pdf * last_pdf + pl * pl);
                        L = add(L, mul(thp, scale(m.emissive, w)));
                    } else {
                        L = add(L, mul(thp, m.emissive));
                    }
                    rad[pixel] = make_float4(L.x, L.y, L.z, L4.w);
                } else {
                    f3 wo = neg(dir);
                    f3 nn = hi.normal;
                    if (dot(nn, wo) < 0.0f) nn = neg(nn);
                    NeeOut ne = nee_candidate(S, hi.pos, nn, wo, m, seed);
                    f3 sorg = add(hi.pos, scale(nn, 1e-4f));
                    if (ne.shadow) {
                        push_sh = true;
                        f3 c = mul(thp, ne.c);
                        so = make_float4(sorg.x, sorg.y, sorg.z, ne.tmax);
                        sd = make_float4(ne.dir.x, ne.dir.y, ne.dir.z, __uint_as_float(pixel));
                        sc = make_float4(c.x, c.y, c.z, 0.0f);
                    }
                    f3 wi; float pdf; f3 f;
                    if (A.guide_mode != RRH_GUIDE_NONE) {
                        f = guided_sample(A.guide, A.guide_mode, A.guide_alpha, S, i, hi.pos, nn, wo, m, seed, wi, pdf);
                    } else {
                        float ux = rnd(seed);
                        float uy = rnd(seed);
                        f = mat_sample(m, wo, nn, ux, uy, wi, pdf);
                    }
                    bool alive = !(pdf < 1e-6f);
                    f3 thp_before = thp;
                    f3 bw = mk(0, 0, 0);
                    if (alive) {
                        float ndw = fmaxf(dot(nn, wi), 0.0f);
                        bw = divs(scale(f, ndw), pdf);
                        thp = mul(thp, bw);
                        if (b >= A.rr_start) {
                            float survive = fminf(fmaxf(thp.x, fmaxf(thp.y, thp.z)), 0.95f);
                            if (rnd(seed) > survive) alive = false;
                            else thp = divs(thp, survive);
                        }
                    }
                    if (A.rec.v) guide_record_vertex(A.rec, pixel, b, hi.pos, nn, wo, wi, pdf, bw,
                                                     alive ? thp : mk(0, 0, 0), L, ne.shadow ? mul(thp_before, ne.c) : mk(0, 0, 0), hi.mat, alive);
                    if (alive && !last) {
                        push_ext = true;
                        no = make_float4(sorg.x, sorg.y, sorg.z, __uint_as_float(pixel));
                        nd = make_float4(wi.x, wi.y, wi.z, __uint_as_float(seed));
                        nt = make_float4(thp.x, thp.y, thp.z, pdf);
                    }
                }
            }
        }
        uint32_t es = block_push(&cnt[b + 1], push_ext, s_cnt);
        if (push_ext) { out.o[es] = no; out.d[es] = nd; out.thp[es] = nt; }
        uint32_t ss = block_push(&cnt[64 + b], push_sh, s_cnt);
        if (push_sh) { sh_o[ss] = so; sh_d[ss] = sd; sh_c[ss] = sc; }
    }
}

// COUNT: node and triangle fetches and rays into fetch[3..5]
template <int STK, bool TOP, bool COUNT>
__global__ void __launch_bounds__(kTraceBlock) k_shadow(DevScene S, const uint32_t* __restrict__ cnt, uint32_t b,
                                                        const float4* __restrict__ sh_o, const float4* __restrict__ sh_d,
                                                        const float4* __restrict__ sh_c, float4* __restrict__ rad,
                                                        VertexRec rec, bool remap, uint32_t top_n,
                                                        unsigned long long* __restrict__ fetch) {
    __shared__ int s_stack[(kTraceBlock / 64) * StackSel<STK>::kLds * 64];
    const auto top = stage_top<TOP>(S, s_top, top_n);
    auto st = make_stack<STK>(s_stack);
    uint32_t nn = 0, nt = 0, nr = 0;
    trace_shadow_part<COUNT>(S, cnt[64 + b], b, xcd_block(blockIdx.x, gridDim.x, remap), gridDim.x, sh_o, sh_d, sh_c, rad,
                             rec, st, top, nn, nt, nr);
    if constexpr (COUNT) {
        atomicAdd(fetch + 3, (unsigned long long)nn);
        atomicAdd(fetch + 4, (unsigned long long)nt);
        atomicAdd(fetch + 5, (unsigned long long)nr);
        wave_count_max(nn, nt, fetch + 8);
    }
}

// The extension rays of bounce b and the shadow rays of bounce b - 1 in one launch: both
// queues depend only on the shading pass before them, and neither kernel's tail (its
// slowest waves) then runs alone on the chip.  Workgroups [0, g_sh) trace the shadow
// queue, the rest the extension queue.
template <int STK, bool PREP = false>
__global__ void __launch_bounds__(kTraceBlock) k_trace_pair(DevScene S, Q in, const uint32_t* __restrict__ cnt, uint32_t b,
                                                            float4* __restrict__ hits, const uint32_t* __restrict__ perm,
                                                            uint32_t g_sh, const float4* __restrict__ sh_o,
                                                            const float4* __restrict__ sh_d, const float4* __restrict__ sh_c,
                                                            float4* __restrict__ rad, VertexRec rec, bool remap,
                                                            float4* __restrict__ prep) {
    __shared__ int s_stack[(kTraceBlock / 64) * StackSel<STK>::kLds * 64];
    auto st = make_stack<STK>(s_stack);
    const TopNone top;
    uint32_t nn = 0, nt = 0, nr = 0;
    if (blockIdx.x < g_sh) {
        trace_shadow_part<false>(S, cnt[64 + b - 1], b - 1, xcd_block(blockIdx.x, g_sh, remap), g_sh, sh_o, sh_d, sh_c,
                                 rad, rec, st, top, nn, nt, nr);
    } else {
        const uint32_t ng = gridDim.x - g_sh;
        trace_ext_part<false, PREP>(S, in, cnt[b], xcd_block(blockIdx.x - g_sh, ng, remap), ng, hits, perm, prep, st, top, nn,
                                    nt, nr);
    }
}

template <int STK>
int launch_pair(rrh_ctx* c, const DevScene& S, Q& qi, uint32_t* cnt, uint32_t b, uint32_t gt, const uint32_t* perm,
                float4* rad, const VertexRec& rec, const Bufs& B, float4* prep = nullptr) {
    if (prep)
        hipLaunchKernelGGL((k_trace_pair<STK, true>), dim3(2 * gt), dim3(kTraceBlock), 0, c->stream, S, qi,
                           (const uint32_t*)cnt, b, B.hits, perm, gt, (const float4*)B.sh_o, (const float4*)B.sh_d,
                           (const float4*)B.sh_c, rad, rec, c->xcd_remap, prep);
    else
        hipLaunchKernelGGL((k_trace_pair<STK, false>), dim3(2 * gt), dim3(kTraceBlock), 0, c->stream, S, qi,
                           (const uint32_t*)cnt, b, B.hits, perm, gt, (const float4*)B.sh_o, (const float4*)B.sh_d,
                           (const float4*)B.sh_c, rad, rec, c->xcd_remap, prep);
    RRH_LAUNCH_CHECK();
    return RRH_OK;
}

// a launch's queue argument that leaves that part of the launch out
constexpr uint32_t kNoQueue = ~0u;

// the path tracer's rays through the resumable traversal (rrh_device.h): tmin 0.001, no culling
template <bool SHADOW, class Stk>
__device__ __forceinline__ bool trace4q_steps(const DevScene& S, const WRay& r, float& t, uint32_t& prim, float& b1,
                                              float& b2, int& node, int& sp, Stk& st, uint32_t budget) {
    return rrh::trace4q_steps<SHADOW>(S, r, mk(0.0f, 0.0f, 0.0f), 0.001f, false, t, prim, b1, b2, node, sp, st, budget);
}
template <class Stk>
__device__ __forceinline__ bool trace4q_steps(bool shadow, const DevScene& S, const WRay& r, float& t, uint32_t& prim,
                                              float& b1, float& b2, int& node, int& sp, Stk& st, uint32_t budget) {
    return shadow ? trace4q_steps<true>(S, r, t, prim, b1, b2, node, sp, st, budget)
                  : trace4q_steps<false>(S, r, t, prim, b1, b2, node, sp, st, budget);
}

// a finished ray's result: an extension ray's hit record, or a visible shadow ray's NEE
// contribution added to its pixel (every pixel has at most one shadow ray per bounce)
__device__ __forceinline__ void finish_ray(bool shadow, uint32_t q, float t, uint32_t prim, float b1, float b2,
                                           float4* __restrict__ hits, const float4* __restrict__ sh_d,
                                           const float4* __restrict__ sh_c, float4* __restrict__ rad,
                                           const VertexRec& rec, uint32_t b_sh) {
    if (!shadow) {
        hits[q] = make_float4(t, __uint_as_float(prim), b1, b2);
    } else if (prim != 0u) {
        const uint32_t pixel = __float_as_uint(sh_d[q].w);
        const float4 cc = sh_c[q];
        float4 L = rad[pixel];
        L.x = L.x + cc.x; L.y = L.y + cc.y; L.z = L.z + cc.z;
        rad[pixel] = L;
        if (rec.v) guide_record_nee_visible(rec, pixel, b_sh);
    }
}

// Persistent traversal over the quantized 4-wide tree.  One launch drains a virtual queue: the
// shadow rays of bounce b_sh (slots [0, n_sh)), then the extension rays of bounce b_ext (slots
// [n_sh, n_sh + n_ext)); kNoQueue leaves a part out.  A wave keeps its lanes busy: whenever at
// least `refill` of them are idle (at the start all 64) it hands them the next slots with one
// atomic, so a lane whose ray ends early takes another instead of idling until the slowest ray
// of its wave is done; every lane advances its ray one outer step at a time.  Hit records and
// images equal those of k_trace_pair bit for bit.
// Exit: `refill` is 1..64, so a wave whose lanes are all idle always asks for slots, and once the
// queue has none left (`drained`, wave-uniform) it ends as soon as its last lane finishes.
template <int STK>
__global__ void __launch_bounds__(kTraceBlock) k_trace_queue_p(DevScene S, Q in, const uint32_t* __restrict__ cnt,
                                                               uint32_t b_ext, uint32_t b_sh, float4* __restrict__ hits,
                                                               const float4* __restrict__ sh_o,
                                                               const float4* __restrict__ sh_d,
                                                               const float4* __restrict__ sh_c, float4* __restrict__ rad,
                                                               VertexRec rec, uint32_t* __restrict__ next,
                                                               uint32_t refill) {
    __shared__ int s_stack[StackSel<STK>::kLds * 64];
    auto st = make_stack<STK>(s_stack);
    const uint32_t n_sh = (b_sh != kNoQueue) ? cnt[64 + b_sh] : 0u;
    const uint32_t n = n_sh + ((b_ext != kNoQueue) ? cnt[b_ext] : 0u);
    WRay r = wray_setup(mk(0, 0, 0), mk(0, 0, 1));
    float t = 0.0f, b1 = 0.0f, b2 = 0.0f;
    uint32_t prim = ~0u, slot = 0u;
    int node = kTravDone, sp = 0;
    bool active = false, drained = false, shadow = false;
    while (true) {
        const bool idle = !active;
        if (!drained && __popcll(__ballot(idle ? 1 : 0)) >= (int)refill) {
            uint32_t my = 0;
            if (wave_refill(next, n, idle, drained, my)) {
                slot = my;
                active = true;
                shadow = my < n_sh;
                float4 o, d;
                if (shadow) {
                    o = sh_o[my];
                    d = sh_d[my];
                    t = o.w;
                } else {
                    o = in.o[my - n_sh];
                    d = in.d[my - n_sh];
                    t = 1e9f;
                }
                r = wray_setup(mk(o.x, o.y, o.z), mk(d.x, d.y, d.z));
                prim = ~0u;
                b1 = 0.0f;
                b2 = 0.0f;
                sp = 0;
                node = (shadow && !(t >= 0.001f)) ? kTravDone : 0;  // an empty extent: never occluded
            }
        }
        if (__ballot(active ? 1 : 0) == 0ull) {
            if (drained) break;
            continue;
        }
        if (active && trace4q_steps(shadow, S, r, t, prim, b1, b2, node, sp, st, 1u)) {
            finish_ray(shadow, shadow ? slot : slot - n_sh, t, prim, b1, b2, hits, sh_d, sh_c, rad, rec, b_sh);
            active = false;
        }
    }
}

// ---- suspend and resume (RRH_OPT_SUSPEND_VISITS) ---------------------------------------------
// The first launch traces one ray per lane as k_trace_pair does, but a ray that has visited
// `budget` nodes stops at the start of an outer step and is appended to a continuation queue with
// where it stands; a second launch resumes those rays, compacted into full waves.  So the first
// launch's waves end at the budget instead of at their slowest ray, while every wave still starts
// on neighbouring, coherent rays.  Extension and shadow rays have queues of their own (records
// [0, cap) and [cap, 2 cap)); a record (`words` uint4) is (slot, node, sp, prim), (t, b1, b2, 0),
// then the sp stack entries.  A lane that finds its queue full keeps tracing.
struct SuspQ {
    uint4* rec;
    uint32_t *count_ext, *count_sh;
    uint32_t cap, words;
};

// first launch: workgroups [0, g_sh) take the shadow rays of bounce b_sh, the rest the extension
// rays of bounce b_ext (the wave-uniform loop keeps a wave's lanes together for the append)
template <bool SHADOW, class Stk>
__device__ __forceinline__ void trace_part_susp(const DevScene& S, const Q& in, uint32_t n, uint32_t blk, uint32_t nblk,
                                                Stk& st, float4* __restrict__ hits, const float4* __restrict__ sh_o,
                                                const float4* __restrict__ sh_d, const float4* __restrict__ sh_c,
                                                float4* __restrict__ rad, const VertexRec& rec, uint32_t b_sh,
                                                const SuspQ& sq, uint32_t budget) {
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t* count = SHADOW ? sq.count_sh : sq.count_ext;
    uint4* recs = sq.rec + (SHADOW ? (size_t)sq.cap * sq.words : 0);
    for (uint32_t base = blk * kTraceBlock; base < n; base += nblk * kTraceBlock) {
        const uint32_t i = base + lane;
        const bool valid = i < n;
        float t = 1e9f, b1 = 0.0f, b2 = 0.0f;
        uint32_t prim = ~0u;
        int node = kTravDone, sp = 0;
        WRay r = wray_setup(mk(0, 0, 0), mk(0, 0, 1));
        if (valid) {
            const float4 o = SHADOW ? sh_o[i] : in.o[i], d = SHADOW ? sh_d[i] : in.d[i];
            r = wray_setup(mk(o.x, o.y, o.z), mk(d.x, d.y, d.z));
            if (SHADOW) t = o.w;
            node = (SHADOW && !(t >= 0.001f)) ? kTravDone : 0;  // an empty extent: never occluded
        }
        bool done = trace4q_steps<SHADOW>(S, r, t, prim, b1, b2, node, sp, st, budget);
        // the suspended lanes' records: one atomic per wave
        const unsigned long long m = __ballot(done ? 0 : 1);
        if (m != 0ull) {
            const int leader = __ffsll((long long)m) - 1;
            uint32_t at = 0u;
            if ((int)lane == leader) at = atomicAdd(count, (uint32_t)__popcll(m));
            at = __shfl(at, leader) + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
            if (!done) {
                if (at < sq.cap) {
                    uint4* w = recs + (size_t)at * sq.words;
                    w[0] = make_uint4(i, (uint32_t)node, (uint32_t)sp, prim);
                    w[1] = make_uint4(__float_as_uint(t), __float_as_uint(b1), __float_as_uint(b2), 0u);
                    uint32_t* ws = reinterpret_cast<uint32_t*>(w + 2);
                    for (int k = 0; k < sp; ++k) ws[k] = (uint32_t)st.get(k);
                } else {
                    done = trace4q_steps<SHADOW>(S, r, t, prim, b1, b2, node, sp, st, ~0u);
                }
            }
        }
        if (valid && done) finish_ray(SHADOW, i, t, prim, b1, b2, hits, sh_d, sh_c, rad, rec, b_sh);
    }
}
template <int STK>
__global__ void __launch_bounds__(kTraceBlock) k_trace_susp(DevScene S, Q in, const uint32_t* __restrict__ cnt,
                                                            uint32_t b_ext, uint32_t b_sh, uint32_t g_sh,
                                                            float4* __restrict__ hits, const float4* __restrict__ sh_o,
                                                            const float4* __restrict__ sh_d,
                                                            const float4* __restrict__ sh_c, float4* __restrict__ rad,
                                                            VertexRec rec, SuspQ sq, uint32_t budget, bool remap) {
    __shared__ int s_stack[StackSel<STK>::kLds * 64];
    auto st = make_stack<STK>(s_stack);
    if (blockIdx.x < g_sh)
        trace_part_susp<true>(S, in, cnt[64 + b_sh], xcd_block(blockIdx.x, g_sh, remap), g_sh, st, hits, sh_o, sh_d,
                              sh_c, rad, rec, b_sh, sq, budget);
    else
        trace_part_susp<false>(S, in, cnt[b_ext], xcd_block(blockIdx.x - g_sh, gridDim.x - g_sh, remap),
                               gridDim.x - g_sh, st, hits, sh_o, sh_d, sh_c, rad, rec, b_sh, sq, budget);
}

// second launch: the suspended rays to the end, one per lane in the records' order; workgroups
// [0, g_sh) the shadow queue, the rest the extension queue
template <bool SHADOW, class Stk>
__device__ __forceinline__ void resume_part(const DevScene& S, const Q& in, uint32_t blk, uint32_t nblk, Stk& st,
                                            float4* __restrict__ hits, const float4* __restrict__ sh_o,
                                            const float4* __restrict__ sh_d, const float4* __restrict__ sh_c,
                                            float4* __restrict__ rad, const VertexRec& rec, uint32_t b_sh,
                                            const SuspQ& sq) {
    const uint32_t n = min(*(SHADOW ? sq.count_sh : sq.count_ext), sq.cap);
    const uint4* recs = sq.rec + (SHADOW ? (size_t)sq.cap * sq.words : 0);
    for (uint32_t j = blk * kTraceBlock + (threadIdx.x & 63u); j < n; j += nblk * kTraceBlock) {
        const uint4* w = recs + (size_t)j * sq.words;
        const uint4 h0 = w[0], h1 = w[1];
        const uint32_t q = h0.x;
        int node = (int)h0.y, sp = (int)h0.z;
        uint32_t prim = h0.w;
        float t = __uint_as_float(h1.x), b1 = __uint_as_float(h1.y), b2 = __uint_as_float(h1.z);
        const uint32_t* ws = reinterpret_cast<const uint32_t*>(w + 2);
        for (int k = 0; k < sp; ++k) st.put(k, (int)ws[k]);
        const float4 o = SHADOW ? sh_o[q] : in.o[q], d = SHADOW ? sh_d[q] : in.d[q];
        const WRay r = wray_setup(mk(o.x, o.y, o.z), mk(d.x, d.y, d.z));
        trace4q_steps<SHADOW>(S, r, t, prim, b1, b2, node, sp, st, ~0u);
        finish_ray(SHADOW, q, t, prim, b1, b2, hits, sh_d, sh_c, rad, rec, b_sh);
    }
}
template <int STK>
__global__ void __launch_bounds__(kTraceBlock) k_trace_resume(DevScene S, Q in, uint32_t b_sh, uint32_t g_sh,
                                                              float4* __restrict__ hits,
                                                              const float4* __restrict__ sh_o,
                                                              const float4* __restrict__ sh_d,
                                                              const float4* __restrict__ sh_c, float4* __restrict__ rad,
                                                              VertexRec rec, SuspQ sq) {
    __shared__ int s_stack[StackSel<STK>::kLds * 64];
    auto st = make_stack<STK>(s_stack);
    if (blockIdx.x < g_sh)
        resume_part<true>(S, in, blockIdx.x, g_sh, st, hits, sh_o, sh_d, sh_c, rad, rec, b_sh, sq);
    else
        resume_part<false>(S, in, blockIdx.x - g_sh, gridDim.x - g_sh, st, hits, sh_o, sh_d, sh_c, rad, rec, b_sh, sq);
}

// one-wave workgroups of k_trace_queue_p<STK> a CU holds at once (cached per stack kind)
template <int STK>
uint32_t queue_p_waves_per_cu() {
    static uint32_t per_cu = 0;
    if (per_cu == 0) {
        int v = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, k_trace_queue_p<STK>, kTraceBlock, 0) != hipSuccess || v < 1)
            v = 1;
        per_cu = (uint32_t)v;
    }
    return per_cu;
}

// the persistent launch of one virtual queue (see k_trace_queue_p); next: its slot counter
// (zeroed with the frame's counters); RRH_OPT_PERSISTENT above 1 caps the waves per CU
template <int STK>
int launch_queue_p(rrh_ctx* c, const DevScene& S, const Q& qi, uint32_t* cnt, uint32_t b_ext, uint32_t b_sh,
                   uint32_t* next, float4* rad, const VertexRec& rec, const Bufs& B) {
    uint32_t w = queue_p_waves_per_cu<STK>();
    if (c->persistent_waves > 1) w = std::min(w, c->persistent_waves);
    const uint32_t g = std::max(w, 1u) * (uint32_t)c->cu_count;
    hipLaunchKernelGGL(k_trace_queue_p<STK>, dim3(g), dim3(kTraceBlock), 0, c->stream, S, qi, (const uint32_t*)cnt,
                       b_ext, b_sh, B.hits, (const float4*)B.sh_o, (const float4*)B.sh_d, (const float4*)B.sh_c, rad,
                       rec, next, c->refill);
    RRH_LAUNCH_CHECK();
    return RRH_OK;
}

// first launch plus resume of the suspend mode: g_sh shadow workgroups (bounce b_sh) and g_ext
// extension workgroups (bounce b_ext); the suspended counts go to counters [128 + b_ext] and
// [192 + b_sh] (zeroed with the frame's; no two launches of a frame share one)
template <int STK>
int launch_susp(rrh_ctx* c, const DevScene& S, const Q& qi, uint32_t* cnt, uint32_t b_ext, uint32_t b_sh,
                uint32_t g_sh, uint32_t g_ext, float4* rad, const VertexRec& rec, const Bufs& B, const SuspQ& sq0) {
    SuspQ sq = sq0;
    sq.count_ext = cnt + 128 + b_ext;
    sq.count_sh = cnt + 192 + b_sh;
    hipLaunchKernelGGL(k_trace_susp<STK>, dim3(g_sh + g_ext), dim3(kTraceBlock), 0, c->stream, S, qi,
                       (const uint32_t*)cnt, b_ext, b_sh, g_sh, B.hits, (const float4*)B.sh_o, (const float4*)B.sh_d,
                       (const float4*)B.sh_c, rad, rec, sq, c->suspend_visits, c->xcd_remap);
    RRH_LAUNCH_CHECK();
    // resume: up to 16 waves per CU for each part that has rays (waves without records end at once)
    const uint32_t gpart = std::max(1u, std::min((uint32_t)c->cu_count * 16u, (sq.cap + kTraceBlock - 1) / kTraceBlock));
    const uint32_t rsh = g_sh ? gpart : 0u, rext = g_ext ? gpart : 0u;
    hipLaunchKernelGGL(k_trace_resume<STK>, dim3(rsh + rext), dim3(kTraceBlock), 0, c->stream, S, qi, b_sh, rsh,
                       B.hits, (const float4*)B.sh_o, (const float4*)B.sh_d, (const float4*)B.sh_c, rad, rec, sq);
    RRH_LAUNCH_CHECK();
    return RRH_OK;
}
int launch_susp_sel(int sk, rrh_ctx* c, const DevScene& S, const Q& qi, uint32_t* cnt, uint32_t b_ext, uint32_t b_sh,
                    uint32_t g_sh, uint32_t g_ext, float4* rad, const VertexRec& rec, const Bufs& B, const SuspQ& sq) {
    switch (sk) {
        case RRH_STACK_LDS32: return launch_susp<RRH_STACK_LDS32>(c, S, qi, cnt, b_ext, b_sh, g_sh, g_ext, rad, rec, B, sq);
        case RRH_STACK_HYBRID16: return launch_susp<RRH_STACK_HYBRID16>(c, S, qi, cnt, b_ext, b_sh, g_sh, g_ext, rad, rec, B, sq);
        case RRH_STACK_HYBRID32: return launch_susp<RRH_STACK_HYBRID32>(c, S, qi, cnt, b_ext, b_sh, g_sh, g_ext, rad, rec, B, sq);
        default: return launch_susp<RRH_STACK_PRIVATE>(c, S, qi, cnt, b_ext, b_sh, g_sh, g_ext, rad, rec, B, sq);
    }
}

int launch_queue_p_sel(int sk, rrh_ctx* c, const DevScene& S, const Q& qi, uint32_t* cnt, uint32_t b_ext,
                       uint32_t b_sh, uint32_t* next, float4* rad, const VertexRec& rec, const Bufs& B) {
    switch (sk) {
        case RRH_STACK_LDS32: return launch_queue_p<RRH_STACK_LDS32>(c, S, qi, cnt, b_ext, b_sh, next, rad, rec, B);
        case RRH_STACK_HYBRID16: return launch_queue_p<RRH_STACK_HYBRID16>(c, S, qi, cnt, b_ext, b_sh, next, rad, rec, B);
        case RRH_STACK_HYBRID32: return launch_queue_p<RRH_STACK_HYBRID32>(c, S, qi, cnt, b_ext, b_sh, next, rad, rec, B);
        default: return launch_queue_p<RRH_STACK_PRIVATE>(c, S, qi, cnt, b_ext, b_sh, next, rad, rec, B);
    }
}

// ray-sort key of each queue slot: direction octant (3 bits), then the origin's Morton code
// (10 bits per axis of the scene box, top 29 bits); slots past the queue sort last
__device__ __forceinline__ uint32_t spread10(uint32_t x) {
    x &= 0x3ffu;
    x = (x | x << 16) & 0x030000ffu;
    x = (x | x << 8) & 0x0300f00fu;
    x = (x | x << 4) & 0x030c30c3u;
    x = (x | x << 2) & 0x09249249u;
    return x;
}
__global__ void __launch_bounds__(kBlock) k_ray_keys(Q q, const uint32_t* __restrict__ cnt, uint32_t b, uint32_t npix,
                                                     float3 bmin, float inv_ext, uint32_t shift,
                                                     uint32_t* __restrict__ keys, uint32_t* __restrict__ slots) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= npix) return;
    uint32_t key = 0xFFFFFFFFu;
    if (i < cnt[b]) {
        const float4 o = q.o[i], d = q.d[i];
        const uint32_t oct = (d.x < 0.0f ? 1u : 0u) | (d.y < 0.0f ? 2u : 0u) | (d.z < 0.0f ? 4u : 0u);
        auto cell = [&](float v, float lo) {
            const float t = (v - lo) * inv_ext * 1024.0f;
            return (t >= 0.0f) ? (uint32_t)fminf(t, 1023.0f) : 0u;  // NaN -> 0
        };
        const uint32_t m = spread10(cell(o.x, bmin.x)) << 2 | spread10(cell(o.y, bmin.y)) << 1 | spread10(cell(o.z, bmin.z));
        key = oct << 29 | m >> 1;
    }
    keys[i] = key >> shift;  // the top bits only, sorted over the full 32-bit range
    slots[i] = i;
}

// blocks the device holds at once for a kernel (persistent grids), cached per instantiation
template <class K>
uint32_t resident_blocks(rrh_ctx* c, K kernel) {
    static uint32_t per_cu = 0;
    if (per_cu == 0) {
        int v = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, kernel, kBlock, 0) != hipSuccess || v < 1) v = 1;
        per_cu = (uint32_t)v;
    }
    return per_cu * (uint32_t)c->cu_count;
}

// prep: also write the learned guide's input records (*prep_used says whether the launched
// kernel did: the plain launch does, the counting, LDS-top and persistent binary-node ones do not)
template <int STK>
int launch_trace(rrh_ctx* c, const DevScene& S, Q& qi, uint32_t* cnt, uint32_t b, uint32_t g, uint32_t gt, bool count,
                 const uint32_t* perm, float4* hp, uint32_t top_n, float4* prep = nullptr, bool* prep_used = nullptr) {
    unsigned long long* fc = (unsigned long long*)c->fetch.p;
    if (prep_used) *prep_used = false;
    if constexpr (STK != RRH_STACK_NONE) {  // the stackless traversal is 4-wide only
    if (c->persistent && S.node_fmt == RRH_NODES_BINARY) {
        uint32_t* next = cnt + 128 + b;
        if (count) {
            const uint32_t gp = std::min(g, resident_blocks(c, k_intersect_p<true, STK>));
            hipLaunchKernelGGL((k_intersect_p<true, STK>), dim3(gp), dim3(kBlock), 0, c->stream, S, qi, cnt, b, hp, next, fc);
        } else {
            const uint32_t gp = std::min(g, resident_blocks(c, k_intersect_p<false, STK>));
            hipLaunchKernelGGL((k_intersect_p<false, STK>), dim3(gp), dim3(kBlock), 0, c->stream, S, qi, cnt, b, hp, next, fc);
        }
        RRH_LAUNCH_CHECK();
        return RRH_OK;
    }
    }
    const size_t lds = (size_t)top_n * 4 * sizeof(uint4);
    if (prep && !count && !top_n) {
        hipLaunchKernelGGL((k_intersect<false, STK, false, true>), dim3(gt), dim3(kTraceBlock), 0, c->stream, S, qi, cnt, b,
                           hp, fc, c->xcd_remap, perm, 0u, prep);
        RRH_LAUNCH_CHECK();
        if (prep_used) *prep_used = true;
        return RRH_OK;
    }
    if (count && top_n)
        hipLaunchKernelGGL((k_intersect<true, STK, true>), dim3(gt), dim3(kTraceBlock), lds, c->stream, S, qi, cnt, b, hp, fc,
                           c->xcd_remap, perm, top_n, prep);
    else if (count)
        hipLaunchKernelGGL((k_intersect<true, STK, false>), dim3(gt), dim3(kTraceBlock), 0, c->stream, S, qi, cnt, b, hp, fc,
                           c->xcd_remap, perm, 0u, prep);
    else if (top_n)
        hipLaunchKernelGGL((k_intersect<false, STK, true>), dim3(gt), dim3(kTraceBlock), lds, c->stream, S, qi, cnt, b, hp, fc,
                           c->xcd_remap, perm, top_n, prep);
    else
        hipLaunchKernelGGL((k_intersect<false, STK, false>), dim3(gt), dim3(kTraceBlock), 0, c->stream, S, qi, cnt, b, hp, fc,
                           c->xcd_remap, perm, 0u, prep);
    RRH_LAUNCH_CHECK();
    return RRH_OK;
}
template <int STK>
int launch_shadow(rrh_ctx* c, const DevScene& S, uint32_t* cnt, uint32_t b, uint32_t g, uint32_t gt, float4* rad,
                  const VertexRec& rec, const Bufs& B, uint32_t top_n, bool count) {
    if constexpr (STK != RRH_STACK_NONE) {
    if (c->persistent && S.node_fmt == RRH_NODES_BINARY) {
        const uint32_t gp = std::min(g, resident_blocks(c, k_shadow_p<STK>));
        hipLaunchKernelGGL(k_shadow_p<STK>, dim3(gp), dim3(kBlock), 0, c->stream, S, (const uint32_t*)cnt, b,
                           (const float4*)B.sh_o, (const float4*)B.sh_d, (const float4*)B.sh_c, rad, rec,
                           cnt + 192 + b);
        RRH_LAUNCH_CHECK();
        return RRH_OK;
    }
    }
    unsigned long long* fc = (unsigned long long*)c->fetch.p;
    const size_t lds = (size_t)top_n * 4 * sizeof(uint4);
    auto go = [&](auto k) {
        hipLaunchKernelGGL(k, dim3(gt), dim3(kTraceBlock), lds, c->stream, S, (const uint32_t*)cnt, b,
                           (const float4*)B.sh_o, (const float4*)B.sh_d, (const float4*)B.sh_c, rad, rec, c->xcd_remap,
                           top_n, fc);
    };
    if (top_n && count) go(k_shadow<STK, true, true>);
    else if (top_n) go(k_shadow<STK, true, false>);
    else if (count) go(k_shadow<STK, false, true>);
    else go(k_shadow<STK, false, false>);
    RRH_LAUNCH_CHECK();
    return RRH_OK;
}

__global__ void k_accumulate(const float4* __restrict__ rad, float4* __restrict__ acc, uint32_t n,
                             uint32_t camera_moved, uint32_t spp_in) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 r = rad[i];
    if (camera_moved != 0u || spp_in == 0u) { acc[i] = r; return; }
    float4 p = acc[i];
    float spp = (float)spp_in;
    float w = 1.0f / (spp + 1.0f);
    float k = spp / (spp + 1.0f);
    acc[i] = make_float4(p.x * k + r.x * w, p.y * k + r.y * w, p.z * k + r.z * w, p.w * k + r.w * w);
}

__global__ void k_accumulate_f64(const float4* __restrict__ rad, double* __restrict__ acc, uint32_t n,
                                 uint32_t camera_moved, uint32_t spp_in) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 r = rad[i];
    double* a = acc + (size_t)i * 4;
    if (camera_moved != 0u || spp_in == 0u) {
        a[0] = r.x; a[1] = r.y; a[2] = r.z; a[3] = r.w;
    } else {
        a[0] += (double)r.x; a[1] += (double)r.y; a[2] += (double)r.z; a[3] += (double)r.w;
    }
}

// stack kind of the ray-queue kernels: RRH_STACK_NONE runs stackless on a 4-wide tree of at
// most 31 levels (the trail's reach), other kinds as stack_for gives them
int queue_stack(const rrh_ctx* c, const rrh_scene* s) {
    const int fmt = s->node_format();
    if (c->stack_kind == RRH_STACK_NONE && fmt != RRH_NODES_BINARY && s->depth4 <= 31) return RRH_STACK_NONE;
    return stack_for(c, s->stack_need());
}

uint32_t grid_for(uint32_t n) {
    uint32_t g = (n + kBlock - 1) / kBlock;
    return g < 1u ? 1u : (g > 8192u ? 8192u : g);  // one path per lane up to 2M paths
}

// the context's frame buffers, allocated on first use and grown with the extent
int context_bufs(rrh_ctx* c, uint32_t npix, Bufs& B) {
    int rc = 0;
    for (int k = 0; k < 2; ++k) {
        if ((rc = c->q_o[k].ensure((size_t)npix * 16))) return rc;
        if ((rc = c->q_d[k].ensure((size_t)npix * 16))) return rc;
        if ((rc = c->q_thp[k].ensure((size_t)npix * 16))) return rc;
    }
    if ((rc = c->hits.ensure((size_t)npix * 16))) return rc;
    if ((rc = c->sh_o.ensure((size_t)npix * 16))) return rc;
    if ((rc = c->sh_d.ensure((size_t)npix * 16))) return rc;
    if ((rc = c->sh_c.ensure((size_t)npix * 16))) return rc;
    if ((rc = c->counters.ensure(256 * 4))) return rc;
    if (c->ray_sort_bits) {
        if ((rc = c->sort_keys.ensure((size_t)npix * 8))) return rc;
        if ((rc = c->sort_perm.ensure((size_t)npix * 8))) return rc;
        size_t tb = 0;
        uint32_t* k = (uint32_t*)c->sort_keys.p;
        uint32_t* v = (uint32_t*)c->sort_perm.p;
        RRH_HIP(rocprim::radix_sort_pairs(nullptr, tb, k, k + npix, v, v + npix, npix, 0, 32, c->stream));
        if ((rc = c->sort_tmp.ensure(tb))) return rc;
    }
    for (int k = 0; k < 2; ++k) B.q[k] = Q{(float4*)c->q_o[k].p, (float4*)c->q_d[k].p, (float4*)c->q_thp[k].p};
    B.hits = (float4*)c->hits.p;
    B.sh_o = (float4*)c->sh_o.p;
    B.sh_d = (float4*)c->sh_d.p;
    B.sh_c = (float4*)c->sh_c.p;
    B.cnt = (uint32_t*)c->counters.p;
    return RRH_OK;
}

}  // namespace

// fixed == nullptr: the context's buffers (allocated on first use, grown with the extent);
// otherwise the caller's buffers, and nothing is allocated: unguided frames only, no ray sort,
// no fetch counting, no profiling events, counters cleared by a kernel
static int rrh_pt_render_impl(rrh_ctx* c, rrh_scene* s, const rrh_pt_params* p, rrh_grid* grid, rrh_mlp* mlp,
                              rrh_image* radiance, const Bufs* fixed = nullptr) {
    RRH_ARG(c && s && p && radiance, "rrh_pt_render: null argument");
    RRH_ARG(s->cam_set, "rrh_pt_render: camera not set (rrh_camera_set)");
    RRH_ARG(radiance->fmt == RRH_FMT_RGBA32F, "rrh_pt_render: radiance must be RGBA32F");
    uint32_t W = s->cam_host.screen_width, H = s->cam_host.screen_height;
    RRH_ARG(radiance->w == W && radiance->h == H, "rrh_pt_render: image extent differs from the camera");
    RRH_ARG(p->max_bounces >= 1 && p->max_bounces <= kMaxBounces, "rrh_pt_render: max_bounces must be in [1, 32]");
    RRH_ARG(p->guide_mode <= RRH_GUIDE_MLP, "rrh_pt_render: bad guide_mode");
    const bool band = p->row_begin != 0 || p->row_end != 0;
    const uint32_t r0 = band ? p->row_begin : 0u, r1 = band ? p->row_end : H;
    RRH_ARG(r0 < r1 && r1 <= H, "rrh_pt_render: row band outside the image (row_begin < row_end <= height)");
    RRH_ARG(!band || (p->guide_mode == RRH_GUIDE_NONE && p->record_vertices == 0),
            "rrh_pt_render: a row band is for unguided frames without vertex recording");
    // every per-ray buffer is sized for the band's pixels
    uint32_t npix = W * (r1 - r0);
    int rc = 0;
    Bufs Bf;
    if (fixed) Bf = *fixed;
    else if ((rc = context_bufs(c, npix, Bf))) return rc;
    ShadeArgs A;
    A.max_bounces = p->max_bounces;
    A.rr_start = p->rr_start;
    A.guide_mode = p->guide_mode;
    A.guide_alpha = p->guide_alpha;
    A.remap = c->xcd_remap;
    if ((rc = guide_prepare(c, s, p, grid, mlp, A.guide, A.rec))) return rc;
    // the learned guide's per-slot input records (written by the traversal epilogue or the input pass)
    if (p->guide_mode == RRH_GUIDE_MLP && (rc = c->mlp_prep.ensure((size_t)npix * 48))) return rc;
    DevScene S = s->dev();
    uint32_t* cnt = Bf.cnt;
    if (fixed) {
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(256), 0, c->stream, cnt, 256u);
        RRH_LAUNCH_CHECK();
    } else {
        RRH_HIP(hipMemsetAsync(cnt, 0, 256 * 4, c->stream));
    }
    float4* rad = (float4*)radiance->p;
    Q* q = Bf.q;
    hipLaunchKernelGGL(k_raygen, dim3((W + 7) / 8, (r1 - r0 + 7) / 8), dim3(64), 0, c->stream, S, q[0], cnt, rad,
                       p->frame_seed, r0, r1 - r0);
    RRH_LAUNCH_CHECK();
    uint32_t g = grid_for(npix);
    const uint32_t gt = std::min((npix + kTraceBlock - 1) / kTraceBlock, kTraceGridCap);
    const int sk = queue_stack(c, s);
    const uint32_t top_n = (s->node_format() == RRH_NODES_WIDE4Q) ? std::min(c->top_nodes, s->top4) : 0u;
    const bool cf = (p->flags & RRH_PT_COUNT_FETCHES) != 0;
    // paired launches: not with fetch counting (which counts the closest-hit kernel alone),
    // the LDS top of the tree, or the persistent binary kernels
    const bool pairs = c->trace_pairs && !cf && top_n == 0 && !(c->persistent && S.node_fmt == RRH_NODES_BINARY);
    // persistent waves over the quantized 4-wide tree: each launch drains its queues (the paired
    // shadow queue first); not with fetch counting, the LDS top of the tree, the stackless
    // traversal or the ray sort
    const bool pq = c->persistent && S.node_fmt == RRH_NODES_WIDE4Q && !cf && top_n == 0 && sk != RRH_STACK_NONE &&
                    c->ray_sort_bits == 0;
    // suspend and resume (the same conditions, the speculative traversal, and buffers of its own)
    const bool sq = !pq && !fixed && c->suspend_visits > 0 && S.node_fmt == RRH_NODES_WIDE4Q && S.speculative &&
                    !cf && top_n == 0 && sk != RRH_STACK_NONE && c->ray_sort_bits == 0;
    SuspQ susp{nullptr, nullptr, nullptr, npix, 2u + (s->stack_need() + 3u) / 4u};
    if (sq) {
        if ((rc = c->susp.ensure((size_t)2u * susp.cap * susp.words * 16u))) return rc;
        susp.rec = (uint4*)c->susp.p;
    }
    // profiling: an event pair around every traversal launch (closest hit, paired, shadow)
    const bool prof = c->profiling && !fixed;
    auto prof_begin = [&]() -> int {
        if (!prof) return RRH_OK;
        if (c->prof_used + 2 > c->prof_events.size()) {
            for (int k = 0; k < 64; ++k) {
                hipEvent_t e;
                RRH_HIP(hipEventCreate(&e));
                c->prof_events.push_back(e);
            }
        }
        RRH_HIP(hipEventRecord(c->prof_events[c->prof_used], c->stream));
        return RRH_OK;
    };
    auto prof_end = [&]() -> int {
        if (!prof) return RRH_OK;
        RRH_HIP(hipEventRecord(c->prof_events[c->prof_used + 1], c->stream));
        c->prof_used += 2;
        return RRH_OK;
    };
    for (uint32_t b = 0; b < p->max_bounces; ++b) {
        Q& qi = q[b & 1];
        Q& qo = q[(b + 1) & 1];
        const uint32_t* perm = nullptr;
        if (c->ray_sort_bits && b > 0) {
            // extension rays: trace in key order (primary rays are already tile-coherent)
            uint32_t* keys = (uint32_t*)c->sort_keys.p;
            uint32_t* slots = (uint32_t*)c->sort_perm.p;
            const float ext = std::max(std::max(s->bmax[0] - s->bmin[0], s->bmax[1] - s->bmin[1]), s->bmax[2] - s->bmin[2]);
            hipLaunchKernelGGL(k_ray_keys, dim3((npix + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, qi,
                               (const uint32_t*)cnt, b, npix, make_float3(s->bmin[0], s->bmin[1], s->bmin[2]),
                               ext > 0.0f ? 1.0f / ext : 0.0f, 32u - c->ray_sort_bits, keys, slots);
            RRH_LAUNCH_CHECK();
            // full bit range: a partial range (begin_bit > 0) returned orders that were not
            // permutations of the slots for queues under rocPRIM's merge-sort size limit
            size_t tb = c->sort_tmp.bytes;
            hipError_t e = rocprim::radix_sort_pairs(c->sort_tmp.p, tb, keys, keys + npix, slots, slots + npix, npix,
                                                     0, 32, c->stream);
            if (e != hipSuccess) { set_error(std::string("ray sort: ") + hipGetErrorString(e)); return RRH_E_HIP; }
            perm = slots + npix;
        }
        // the direction sampled at the last bounce is never traced (and its training record
        // carries no radiance): the guide is skipped there
        const bool guided_b = !(b + 1 >= p->max_bounces || (p->guide_bounces != 0 && b >= p->guide_bounces));
        // an MLP-guided bounce: the traversal kernels write the network's input records in their
        // epilogue (the persistent and suspend variants leave that to the input pass)
        float4* prep = (A.guide_mode == RRH_GUIDE_MLP && guided_b && !pq && !sq) ? (float4*)c->mlp_prep.p : nullptr;
        bool prep_done = false;
        if ((rc = prof_begin())) return rc;
        if (pq)
            rc = launch_queue_p_sel(sk, c, S, qi, cnt, b, (pairs && b > 0) ? b - 1 : kNoQueue, cnt + 128 + b, rad,
                                    A.rec, Bf);
        else if (sq)
            rc = launch_susp_sel(sk, c, S, qi, cnt, b, (pairs && b > 0) ? b - 1 : 0u, (pairs && b > 0) ? gt : 0u, gt,
                                 rad, A.rec, Bf, susp);
        else if (pairs && b > 0) {  // this bounce's extension rays with the previous bounce's shadow rays
            rc = (sk == RRH_STACK_LDS32) ? launch_pair<RRH_STACK_LDS32>(c, S, qi, cnt, b, gt, perm, rad, A.rec, Bf, prep)
               : (sk == RRH_STACK_HYBRID16) ? launch_pair<RRH_STACK_HYBRID16>(c, S, qi, cnt, b, gt, perm, rad, A.rec, Bf, prep)
               : (sk == RRH_STACK_HYBRID32) ? launch_pair<RRH_STACK_HYBRID32>(c, S, qi, cnt, b, gt, perm, rad, A.rec, Bf, prep)
               : (sk == RRH_STACK_NONE) ? launch_pair<RRH_STACK_NONE>(c, S, qi, cnt, b, gt, perm, rad, A.rec, Bf, prep)
               : launch_pair<RRH_STACK_PRIVATE>(c, S, qi, cnt, b, gt, perm, rad, A.rec, Bf, prep);
            prep_done = prep != nullptr;
        } else {
            bool* pu = &prep_done;
            rc = (sk == RRH_STACK_LDS32) ? launch_trace<RRH_STACK_LDS32>(c, S, qi, cnt, b, g, gt, cf, perm, Bf.hits, top_n, prep, pu)
               : (sk == RRH_STACK_HYBRID16) ? launch_trace<RRH_STACK_HYBRID16>(c, S, qi, cnt, b, g, gt, cf, perm, Bf.hits, top_n, prep, pu)
               : (sk == RRH_STACK_HYBRID32) ? launch_trace<RRH_STACK_HYBRID32>(c, S, qi, cnt, b, g, gt, cf, perm, Bf.hits, top_n, prep, pu)
               : (sk == RRH_STACK_NONE) ? launch_trace<RRH_STACK_NONE>(c, S, qi, cnt, b, g, gt, cf, perm, Bf.hits, top_n, prep, pu)
               : launch_trace<RRH_STACK_PRIVATE>(c, S, qi, cnt, b, g, gt, cf, perm, Bf.hits, top_n, prep, pu);
        }
        if (rc) return rc;
        RRH_LAUNCH_CHECK();
        if ((rc = prof_end())) return rc;
        A.b = b;
        ShadeArgs Ab = A;
        if (!guided_b) Ab.guide_mode = RRH_GUIDE_NONE;
        if (Ab.guide_mode == RRH_GUIDE_MLP) {
            rc = mlp_infer_queue(c, mlp, S, qi.o, qi.d, (const float4*)Bf.hits, cnt, b, npix,
                                 const_cast<uint16_t*>(A.guide.mlp_lobes), prep_done);
            if (rc) return rc;
            Ab.hit_prep = (const float4*)c->mlp_prep.p;  // the hit lookup the network's inputs took
        }
        hipLaunchKernelGGL(k_shade, dim3(g), dim3(kBlock), 0, c->stream, S, qi, qo, (const float4*)Bf.hits, cnt,
                           Bf.sh_o, Bf.sh_d, Bf.sh_c, rad, Ab);
        RRH_LAUNCH_CHECK();
        if (pairs && b + 1 < p->max_bounces) continue;  // traced with the next bounce's extension rays
        if ((rc = prof_begin())) return rc;
        rc = pq ? launch_queue_p_sel(sk, c, S, qi, cnt, kNoQueue, b, cnt + 192 + b, rad, A.rec, Bf)
           : sq ? launch_susp_sel(sk, c, S, qi, cnt, b, b, gt, 0u, rad, A.rec, Bf, susp)
           : (sk == RRH_STACK_LDS32) ? launch_shadow<RRH_STACK_LDS32>(c, S, cnt, b, g, gt, rad, A.rec, Bf, top_n, cf)
           : (sk == RRH_STACK_HYBRID16) ? launch_shadow<RRH_STACK_HYBRID16>(c, S, cnt, b, g, gt, rad, A.rec, Bf, top_n, cf)
           : (sk == RRH_STACK_HYBRID32) ? launch_shadow<RRH_STACK_HYBRID32>(c, S, cnt, b, g, gt, rad, A.rec, Bf, top_n, cf)
           : (sk == RRH_STACK_NONE) ? launch_shadow<RRH_STACK_NONE>(c, S, cnt, b, g, gt, rad, A.rec, Bf, top_n, cf)
           : launch_shadow<RRH_STACK_PRIVATE>(c, S, cnt, b, g, gt, rad, A.rec, Bf, top_n, cf);
        if (rc) return rc;
        if ((rc = prof_end())) return rc;
    }
    hipLaunchKernelGGL(k_sum_counts, dim3(1), dim3(64), 0, c->stream, (const uint32_t*)cnt, p->max_bounces,
                       (unsigned long long*)c->totals.p);
    RRH_LAUNCH_CHECK();
    c->last_max_bounces = p->max_bounces;
    c->last_npix = npix;
    c->last_counters = cnt;
    return guide_finish(c, s, p, grid, mlp, A.rec, rad);
}

namespace rrh {

// resolves this device's copy of the no-alloc scratch array once per device (the device must be
// current); called by rrh_ctx_create
int pt_scratch_init(int device, void** out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) {
        set_error("rrh_ctx_create: device index above the no-alloc scratch table");
        return RRH_E_RANGE;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[device]) {
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return RRH_E_HIP;
        }
        g_scratch_addr[device] = ptr;
    }
    *out = g_scratch_addr[device];
    return RRH_OK;
}

}  // namespace rrh

extern "C" {

int rrh_pt_render_noalloc(rrh_ctx* c, rrh_scene* s, const rrh_pt_params* p, rrh_image* radiance) {
    RRH_ARG(c && s && p && radiance, "rrh_pt_render_noalloc: null argument");
    RRH_ARG(c->noalloc_scratch, "rrh_pt_render_noalloc: the context has no scratch address");
    RRH_ARG(s->cam_set, "rrh_pt_render_noalloc: camera not set (rrh_camera_set)");
    RRH_ARG(p->guide_mode == RRH_GUIDE_NONE && p->record_vertices == 0,
            "rrh_pt_render_noalloc: unguided frames only (guide_mode NONE, record_vertices 0)");
    RRH_ARG(p->flags == 0, "rrh_pt_render_noalloc: flags must be 0 (no fetch counting)");
    RRH_ARG(c->ray_sort_bits == 0, "rrh_pt_render_noalloc: ray sorting (RRH_OPT_RAY_SORT) needs allocated keys");
    const uint64_t npix = (uint64_t)s->cam_host.screen_width * s->cam_host.screen_height;
    if (npix > kNoAllocMaxPixels) {
        set_error("rrh_pt_render_noalloc: " + std::to_string(s->cam_host.screen_width) + "x" +
                  std::to_string(s->cam_host.screen_height) + " is above the no-alloc limit of " +
                  std::to_string(kNoAllocMaxPixels) + " pixels (512x512)");
        return RRH_E_RANGE;
    }
    // the planes are carved from the array for the largest extent, so their offsets never change
    unsigned char* base = (unsigned char*)c->noalloc_scratch;
    float4* plane[10];
    for (int k = 0; k < 10; ++k) plane[k] = (float4*)(base + (size_t)k * kNoAllocPlane);
    Bufs B;
    B.q[0] = Q{plane[0], plane[1], plane[2]};
    B.q[1] = Q{plane[3], plane[4], plane[5]};
    B.hits = plane[6];
    B.sh_o = plane[7];
    B.sh_d = plane[8];
    B.sh_c = plane[9];
    B.cnt = (uint32_t*)(base + 10u * kNoAllocPlane);
    return rrh_pt_render_impl(c, s, p, nullptr, nullptr, radiance, &B);
}

int rrh_pt_render(rrh_ctx* c, rrh_scene* s, const rrh_pt_params* p, rrh_image* radiance) {
    RRH_ARG(p, "rrh_pt_render: null params");
    RRH_ARG(p->guide_mode == RRH_GUIDE_NONE, "rrh_pt_render: guided modes go through rrh_pt_render_guided");
    return rrh_pt_render_impl(c, s, p, nullptr, nullptr, radiance);
}

int rrh_pt_render_guided(rrh_ctx* c, rrh_scene* s, const rrh_pt_params* p, rrh_grid* grid, rrh_mlp* mlp,
                         rrh_image* radiance) {
    return rrh_pt_render_impl(c, s, p, grid, mlp, radiance);
}

int rrh_pt_ray_counts(rrh_ctx* c, rrh_ray_counts* out) {
    RRH_ARG(c && out, "rrh_pt_ray_counts: null argument");
    out->primary = out->extension = out->shadow = 0;
    if (!c->last_counters || c->last_max_bounces == 0) return RRH_OK;
    uint32_t h[128];
    RRH_HIP(hipMemcpyAsync(h, c->last_counters, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    RRH_HIP(hipStreamSynchronize(c->stream));
    out->primary = h[0];
    for (uint32_t b = 1; b < c->last_max_bounces; ++b) out->extension += h[b];
    for (uint32_t b = 0; b < c->last_max_bounces; ++b) out->shadow += h[64 + b];
    return RRH_OK;
}

int rrh_ray_totals(rrh_ctx* c, rrh_ray_counts* out, int reset) {
    RRH_ARG(c && out, "rrh_ray_totals: null argument");
    out->primary = out->extension = out->shadow = 0;
    unsigned long long h[3];
    RRH_HIP(hipMemcpyAsync(h, c->totals.p, 24, hipMemcpyDeviceToHost, c->stream));
    if (reset) RRH_HIP(hipMemsetAsync(c->totals.p, 0, 24, c->stream));
    RRH_HIP(hipStreamSynchronize(c->stream));
    out->primary = h[0]; out->extension = h[1]; out->shadow = h[2];
    return RRH_OK;
}

int rrh_fetch_counts(rrh_ctx* c, uint64_t* nodes, uint64_t* tris, uint64_t* rays, int reset) {
    RRH_ARG(c && nodes && tris && rays, "rrh_fetch_counts: null argument");
    *nodes = *tris = *rays = 0;
    if (!c->fetch.p) return RRH_OK;
    unsigned long long h[3];
    RRH_HIP(hipMemcpyAsync(h, c->fetch.p, 24, hipMemcpyDeviceToHost, c->stream));
    if (reset) RRH_HIP(hipMemsetAsync(c->fetch.p, 0, 24, c->stream));
    RRH_HIP(hipStreamSynchronize(c->stream));
    *nodes = h[0]; *tris = h[1]; *rays = h[2];
    return RRH_OK;
}

int rrh_fetch_counts_shadow(rrh_ctx* c, uint64_t* nodes, uint64_t* tris, uint64_t* rays, int reset) {
    RRH_ARG(c && nodes && tris && rays, "rrh_fetch_counts_shadow: null argument");
    *nodes = *tris = *rays = 0;
    if (!c->fetch.p) return RRH_OK;
    unsigned long long h[3];
    RRH_HIP(hipMemcpyAsync(h, (unsigned long long*)c->fetch.p + 3, 24, hipMemcpyDeviceToHost, c->stream));
    if (reset) RRH_HIP(hipMemsetAsync((unsigned long long*)c->fetch.p + 3, 0, 24, c->stream));
    RRH_HIP(hipStreamSynchronize(c->stream));
    *nodes = h[0]; *tris = h[1]; *rays = h[2];
    return RRH_OK;
}

int rrh_fetch_wave_max(rrh_ctx* c, uint64_t* closest_nodes, uint64_t* closest_tris, uint64_t* shadow_nodes,
                       uint64_t* shadow_tris, int reset) {
    RRH_ARG(c && closest_nodes && closest_tris && shadow_nodes && shadow_tris, "rrh_fetch_wave_max: null argument");
    *closest_nodes = *closest_tris = *shadow_nodes = *shadow_tris = 0;
    if (!c->fetch.p) return RRH_OK;
    unsigned long long h[4];
    RRH_HIP(hipMemcpyAsync(h, (unsigned long long*)c->fetch.p + 6, 32, hipMemcpyDeviceToHost, c->stream));
    if (reset) RRH_HIP(hipMemsetAsync((unsigned long long*)c->fetch.p + 6, 0, 32, c->stream));
    RRH_HIP(hipStreamSynchronize(c->stream));
    *closest_nodes = h[0]; *closest_tris = h[1]; *shadow_nodes = h[2]; *shadow_tris = h[3];
    return RRH_OK;
}

int rrh_pt_traversal_rays(rrh_ctx* c, uint64_t* closest, uint64_t* shadow, int reset) {
    RRH_ARG(c && closest && shadow, "rrh_pt_traversal_rays: null argument");
    unsigned long long h[2];
    RRH_HIP(hipMemcpyAsync(h, (unsigned long long*)c->totals.p + 3, 16, hipMemcpyDeviceToHost, c->stream));
    if (reset) RRH_HIP(hipMemsetAsync((unsigned long long*)c->totals.p + 3, 0, 16, c->stream));
    RRH_HIP(hipStreamSynchronize(c->stream));
    *closest = h[0]; *shadow = h[1];
    return RRH_OK;
}

int rrh_profile_enable(rrh_ctx* c, int on) {
    RRH_ARG(c, "rrh_profile_enable: null context");
    c->profiling = on != 0;
    return RRH_OK;
}

int rrh_profile_traversal(rrh_ctx* c, float* total_ms, uint32_t* launches, int reset) {
    RRH_ARG(c && total_ms && launches, "rrh_profile_traversal: null argument");
    *total_ms = 0.0f;
    *launches = c->prof_used / 2;
    if (c->prof_used) RRH_HIP(hipEventSynchronize(c->prof_events[c->prof_used - 1]));
    double acc = 0.0;
    for (uint32_t k = 0; k + 1 < c->prof_used; k += 2) {
        float ms = 0.0f;
        RRH_HIP(hipEventElapsedTime(&ms, c->prof_events[k], c->prof_events[k + 1]));
        acc += ms;
    }
    *total_ms = (float)acc;
    if (reset) c->prof_used = 0;
    return RRH_OK;
}

int rrh_profile_mark(rrh_ctx* c) {
    RRH_ARG(c, "rrh_profile_mark: null context");
    if (!c->prof_mark) RRH_HIP(hipEventCreate(&c->prof_mark));
    RRH_HIP(hipEventRecord(c->prof_mark, c->stream));
    return RRH_OK;
}

// [begin, end] of every profiled traversal launch of `c` in ms after `origin`'s mark; events of
// one device share a clock, so spans of contexts on different streams compare directly
int rrh_profile_spans(rrh_ctx* c, const rrh_ctx* origin, float* begin_ms, float* end_ms, uint32_t cap,
                      uint32_t* n) {
    RRH_ARG(c && origin && begin_ms && end_ms && n, "rrh_profile_spans: null argument");
    RRH_ARG(origin->prof_mark, "rrh_profile_spans: the origin context has no mark (rrh_profile_mark)");
    RRH_ARG(origin->device == c->device, "rrh_profile_spans: contexts on different devices");
    const uint32_t k = c->prof_used / 2;
    *n = k;
    RRH_ARG(cap >= k, "rrh_profile_spans: fewer slots than profiled launches");
    if (c->prof_used) RRH_HIP(hipEventSynchronize(c->prof_events[c->prof_used - 1]));
    RRH_HIP(hipEventSynchronize(origin->prof_mark));
    for (uint32_t i = 0; i < k; ++i) {
        RRH_HIP(hipEventElapsedTime(&begin_ms[i], origin->prof_mark, c->prof_events[2 * i]));
        RRH_HIP(hipEventElapsedTime(&end_ms[i], origin->prof_mark, c->prof_events[2 * i + 1]));
    }
    return RRH_OK;
}

int rrh_accumulate(rrh_ctx* c, const rrh_image* radiance, rrh_image* acc, uint32_t camera_moved, uint32_t spp) {
    RRH_ARG(c && radiance && acc, "rrh_accumulate: null argument");
    RRH_ARG(radiance->fmt == RRH_FMT_RGBA32F && acc->fmt == RRH_FMT_RGBA32F, "rrh_accumulate: RGBA32F images required");
    RRH_ARG(radiance->w == acc->w && radiance->h == acc->h, "rrh_accumulate: extent mismatch");
    uint32_t n = radiance->w * radiance->h;
    hipLaunchKernelGGL(k_accumulate, dim3((n + 255) / 256), dim3(256), 0, c->stream, (const float4*)radiance->p,
                       (float4*)acc->p, n, camera_moved, spp);
    RRH_LAUNCH_CHECK();
    return RRH_OK;
}

int rrh_accumulate_rows(rrh_ctx* c, const rrh_image* radiance, rrh_image* acc, uint32_t camera_moved, uint32_t spp,
                        uint32_t row_begin, uint32_t row_end) {
    RRH_ARG(c && radiance && acc, "rrh_accumulate_rows: null argument");
    RRH_ARG(radiance->fmt == RRH_FMT_RGBA32F && acc->fmt == RRH_FMT_RGBA32F, "rrh_accumulate_rows: RGBA32F images required");
    RRH_ARG(radiance->w == acc->w && radiance->h == acc->h, "rrh_accumulate_rows: extent mismatch");
    RRH_ARG(row_begin < row_end && row_end <= radiance->h, "rrh_accumulate_rows: rows outside the image (row_begin < row_end <= height)");
    const size_t first = (size_t)row_begin * radiance->w;
    const uint32_t n = (row_end - row_begin) * radiance->w;
    hipLaunchKernelGGL(k_accumulate, dim3((n + 255) / 256), dim3(256), 0, c->stream,
                       (const float4*)radiance->p + first, (float4*)acc->p + first, n, camera_moved, spp);
    RRH_LAUNCH_CHECK();
    return RRH_OK;
}

int rrh_accumulate_f64(rrh_ctx* c, const rrh_image* radiance, rrh_image* acc, uint32_t camera_moved, uint32_t spp) {
    RRH_ARG(c && radiance && acc, "rrh_accumulate_f64: null argument");
    RRH_ARG(radiance->fmt == RRH_FMT_RGBA32F && acc->fmt == RRH_FMT_RGBA64F, "rrh_accumulate_f64: RGBA32F -> RGBA64F");
    RRH_ARG(radiance->w == acc->w && radiance->h == acc->h, "rrh_accumulate_f64: extent mismatch");
    uint32_t n = radiance->w * radiance->h;
    hipLaunchKernelGGL(k_accumulate_f64, dim3((n + 255) / 256), dim3(256), 0, c->stream, (const float4*)radiance->p,
                       (double*)acc->p, n, camera_moved, spp);
    RRH_LAUNCH_CHECK();
    return RRH_OK;
}

}  // extern "C"

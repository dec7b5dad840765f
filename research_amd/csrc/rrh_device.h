// rrh_device.h -- device math shared by the gfx950 kernels.
//
// The shading arithmetic follows the reference Slang shaders operation by
// operation (fp32, evaluated in source order; the library is compiled with
// -ffp-contract=off so no FMA contraction changes a rounding).  sin/cos/exp/
// log are the project's own routines (Cody-Waite reduction + Taylor/atanh
// series) so results do not depend on a vendor math library's ULP behaviour.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "../../include/rrh.h"

#define RRH_D __device__ __forceinline__

namespace rrh {

struct f3 { float x, y, z; };
__host__ __device__ __forceinline__ f3 mk(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
RRH_D f3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }
RRH_D f3 xyz(float4 a) { return mk(a.x, a.y, a.z); }
RRH_D f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
RRH_D f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
RRH_D f3 mul(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
RRH_D f3 scale(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
RRH_D f3 divs(f3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
RRH_D f3 neg(f3 a) { return mk(-a.x, -a.y, -a.z); }
RRH_D float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RRH_D f3 cross(f3 a, f3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
RRH_D float length(f3 a) { return sqrtf(dot(a, a)); }
RRH_D f3 normalize(f3 a) { float inv = 1.0f / sqrtf(dot(a, a)); return scale(a, inv); }
RRH_D float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
RRH_D float lerpf(float x, float y, float s) { return x * (1.0f - s) + y * s; }
RRH_D float mrow(const float* M, int r, float x, float y, float z, float w) {
    return M[r * 4 + 0] * x + M[r * 4 + 1] * y + M[r * 4 + 2] * z + M[r * 4 + 3] * w;
}
RRH_D float comp(f3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

// ---------------------------------------------------------------------------
// deterministic transcendentals
// ---------------------------------------------------------------------------
RRH_D void sincos_rr(float x, float* s, float* c) {
    float k = floorf(x * 0.636619747f + 0.5f);
    float r = x - k * 1.5703125f;
    r = r - k * 4.837512969970703e-4f;
    r = r - k * 7.549790126404332e-8f;
    int q = (int)k;
    float r2 = r * r;
    float sp = r + r * r2 * (-0.166666672f + r2 * (8.33333377e-3f + r2 * (-1.98412701e-4f + r2 * 2.75573188e-6f)));
    float cp = 1.0f + r2 * (-0.5f + r2 * (4.16666679e-2f + r2 * (-1.38888892e-3f + r2 * (2.48015876e-5f + r2 * -2.75573188e-7f))));
    int qq = q & 3;
    float ss = (qq == 0) ? sp : ((qq == 1) ? cp : ((qq == 2) ? -sp : -cp));
    float cc = (qq == 0) ? cp : ((qq == 1) ? -sp : ((qq == 2) ? -cp : sp));
    *s = ss; *c = cc;
}
RRH_D float exp_rr(float x) {
    if (x < -87.0f) return 0.0f;
    if (x > 88.0f) x = 88.0f;
    float k = floorf(x * 1.44269502f + 0.5f);
    float r = x - k * 0.693145751953125f;
    r = r - k * 1.42860677e-6f;
    float p = 1.0f + r * (1.0f + r * (0.5f + r * (0.166666672f + r * (4.16666679e-2f + r * (8.33333377e-3f + r * (1.38888892e-3f + r * 1.98412701e-4f))))));
    int ki = (int)k;
    return p * __uint_as_float((uint32_t)(ki + 127) << 23);
}
// exp(x) of the A-trous config mode's weights (x <= 0 there; no reference counterpart, so the
// arithmetic is this project's own): x = k ln2 + r by a two-part fused reduction, e^r by a
// degree-6 Taylor polynomial in fused multiply-adds (|r| <= ln2 / 2: relative error about 2e-7),
// times 2^k.  Every step is one correctly rounded operation, so the CPU restatement (fmaf) gives
// the same bits.
RRH_D float wexp_rr(float x) {
    if (x < -87.0f) return 0.0f;
    if (x > 88.0f) x = 88.0f;
    const float k = floorf(__builtin_fmaf(x, 1.44269502f, 0.5f));
    float r = __builtin_fmaf(-k, 0.693145751953125f, x);
    r = __builtin_fmaf(-k, 1.42860677e-6f, r);
    float p = __builtin_fmaf(r, 1.38888892e-3f, 8.33333377e-3f);
    p = __builtin_fmaf(r, p, 4.16666679e-2f);
    p = __builtin_fmaf(r, p, 0.166666672f);
    p = __builtin_fmaf(r, p, 0.5f);
    p = __builtin_fmaf(r, p, 1.0f);
    p = __builtin_fmaf(r, p, 1.0f);
    return p * __uint_as_float((uint32_t)((int)k + 127) << 23);
}
RRH_D float log_rr(float x) {
    if (!(x > 0.0f)) return (x == 0.0f) ? -__builtin_inff() : __builtin_nanf("");
    if (__builtin_isinf(x)) return x;
    int e = 0;
    uint32_t u = __float_as_uint(x);
    if (u < 0x00800000u) { u = __float_as_uint(x * 8388608.0f); e = -23; }
    e += (int)((u >> 23) & 255u) - 127;
    float m = __uint_as_float((u & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421354f) { m = m * 0.5f; e = e + 1; }
    float s = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float p = s * (2.0f + s2 * (0.666666687f + s2 * (0.400000006f + s2 * (0.285714298f + s2 * 0.222222224f))));
    float ef = (float)e;
    return ef * 0.693145751953125f + (p + ef * 1.42860677e-6f);
}

// ---------------------------------------------------------------------------
// RNG: Common.slang:18-28, Sampling.slang:11-30
// ---------------------------------------------------------------------------
RRH_D uint32_t pcg_hash(uint32_t seed) {
    uint32_t state = seed * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
RRH_D float u01(uint32_t v) { return (float)(v >> 8) * (1.0f / 16777216.0f); }
RRH_D float rnd(uint32_t& seed) { seed = pcg_hash(seed); return u01(seed); }
RRH_D uint32_t make_seed(uint32_t px, uint32_t py, uint32_t frame) {
    return pcg_hash(px ^ (py * 1000003u) ^ (frame * 2654435761u));
}

// ---------------------------------------------------------------------------
// camera rays: pathtracer.slang:195-205; pixel-centre sky ray restir_di.slang:355-358
// ---------------------------------------------------------------------------
RRH_D f3 cam_dir(const rrh_camera& cam, float ndcx, float ndcy) {
    float vx = mrow(cam.inv_proj, 0, ndcx, ndcy, 1.0f, 1.0f);
    float vy = mrow(cam.inv_proj, 1, ndcx, ndcy, 1.0f, 1.0f);
    f3 w = mk(mrow(cam.inv_view, 0, vx, vy, -1.0f, 0.0f), mrow(cam.inv_view, 1, vx, vy, -1.0f, 0.0f),
              mrow(cam.inv_view, 2, vx, vy, -1.0f, 0.0f));
    return normalize(w);
}
RRH_D f3 primary_dir(const rrh_camera& cam, uint32_t px, uint32_t py, uint32_t& seed) {
    float u0 = rnd(seed);
    float u1 = rnd(seed);
    float ndcx = ((float)px + u0) / (float)cam.screen_width * 2.0f - 1.0f;
    float ndcy = ((float)py + u1) / (float)cam.screen_height * 2.0f - 1.0f;
    return cam_dir(cam, ndcx, ndcy);
}
RRH_D f3 center_dir(const rrh_camera& cam, uint32_t px, uint32_t py) {
    float ndcx = ((float)px + 0.5f) / (float)cam.screen_width * 2.0f - 1.0f;
    float ndcy = ((float)py + 0.5f) / (float)cam.screen_height * 2.0f - 1.0f;
    return cam_dir(cam, ndcx, ndcy);
}
RRH_D f3 sky(f3 d) {
    float t = 0.5f * (d.y + 1.0f);
    return mk(lerpf(0.1f, 0.2f, t), lerpf(0.1f, 0.3f, t), lerpf(0.1f, 0.5f, t));
}

// ---------------------------------------------------------------------------
// device scene
// ---------------------------------------------------------------------------
struct DevScene {
    const float4* nodes;       // 4 float4 per binary node (child boxes + child links)
    const float4* nodes4;      // 8 float4 per 4-wide node (the same tree collapsed, same leaves)
    const uint4* nodesq;       // 4 uint4 per quantized 4-wide node (same links as nodes4)
    const uint32_t* parent4;   // parent of each 4-wide node (stackless traversal)
    uint32_t node_fmt;         // RRH_NODES_*: which of the three the traversal reads
    uint32_t speculative;      // 1: traversal postpones leaves (Aila and Laine 2009; binary and quantized 4-wide)
    const float4* tris;        // 3 float4 per triangle, leaf order; tris[3i].w = flattened id bits
    const uint2* tri_ref;      // per flattened id: (instance, primitive within mesh)
    const uint4* tri_vtx;      // per flattened id: (vertex_base + idx[3 p + k] for k < 3, instance)
    const uint32_t* tri_light; // per flattened id: light index or ~0u
    const float4* tl;          // per triangle light, kTlF4 float4: (v0, Le.x) (v1, Le.y) (v2, Le.z)
                               // (id bits, area, 0, 0) (unit normal, 0)
    const rrh_vertex* verts;
    const uint32_t* idx;
    const rrh_mesh* meshes;
    const rrh_material* mats;
    const rrh_instance* insts;
    const rrh_light* lights;
    const rrh_camera* cam;
    uint32_t nl, ntl, ntri, flags;
};

// ---------------------------------------------------------------------------
// watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2013), two-sided
// ---------------------------------------------------------------------------
struct WRay { int kx, ky, kz; float Sx, Sy, Sz; f3 o, inv; };

RRH_D WRay wray_setup(f3 o, f3 d) {
    WRay r;
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = (ax > ay) ? ((ax > az) ? 0 : 2) : ((ay > az) ? 1 : 2);
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    float dkz = comp(d, kz);
    if (dkz < 0.0f) { int t = kx; kx = ky; ky = t; }
    r.kx = kx; r.ky = ky; r.kz = kz;
    r.Sx = comp(d, kx) / dkz;
    r.Sy = comp(d, ky) / dkz;
    r.Sz = 1.0f / dkz;
    r.o = o;
    r.inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    return r;
}

RRH_D bool wt_hit(const WRay& r, f3 v0, f3 v1, f3 v2, float tmin, float tmax, float& t, float& b1, float& b2) {
    f3 A = sub(v0, r.o), B = sub(v1, r.o), C = sub(v2, r.o);
    float Az = comp(A, r.kz), Bz = comp(B, r.kz), Cz = comp(C, r.kz);
    float Ax = comp(A, r.kx) - r.Sx * Az, Ay = comp(A, r.ky) - r.Sy * Az;
    float Bx = comp(B, r.kx) - r.Sx * Bz, By = comp(B, r.ky) - r.Sy * Bz;
    float Cx = comp(C, r.kx) - r.Sx * Cz, Cy = comp(C, r.ky) - r.Sy * Cz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    Az = r.Sz * Az; Bz = r.Sz * Bz; Cz = r.Sz * Cz;
    float T = U * Az + V * Bz + W * Cz;
    float rcp = 1.0f / det;
    float tt = T * rcp;
    if (!(tt >= tmin && tt <= tmax)) return false;
    t = tt; b1 = V * rcp; b2 = W * rcp;
    return true;
}

// conservative slab test (far distance scaled by 1 + 2*gamma(3), PBRT-v3 3.9.2)
RRH_D bool box_hit(float lox, float hix, float loy, float hiy, float loz, float hiz, const WRay& r,
                   float tmin, float tmax, float& tnear) {
    float t0x = (lox - r.o.x) * r.inv.x, t1x = (hix - r.o.x) * r.inv.x;
    float t0y = (loy - r.o.y) * r.inv.y, t1y = (hiy - r.o.y) * r.inv.y;
    float t0z = (loz - r.o.z) * r.inv.z, t1z = (hiz - r.o.z) * r.inv.z;
    float tn = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), tmin));
    // tmax is the current closest t: widened as well so an exact tie (equal t, smaller id) is never culled
    float tf = fminf(fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z)) * 1.0000008f, tmax * 1.000004f);
    tnear = tn;
    return tn <= tf;
}

RRH_D bool better(float t, uint32_t prim, float bt, uint32_t bp) { return t < bt || (t == bt && prim < bp); }

// child link encoding: >= 0 inner node index; < 0 leaf = ~((first << 3) | (count - 1))
#define RRH_LEAF_FIRST(c) ((uint32_t)(~(c)) >> 3)
#define RRH_LEAF_COUNT(c) ((((uint32_t)(~(c))) & 7u) + 1u)
#define RRH_STACK 64       // stack entries: any BVH the builder accepts (depth <= 63)
#define RRH_HYB_LDS 16     // hybrid stack: entries kept in LDS, deeper ones in a private overflow array

struct Hit { float t; uint32_t prim; float b1, b2; };

// traversal stacks: a lane's entries in LDS are 64 words apart (one wave's lanes
// hit 64 consecutive words, so there are no bank conflicts); the private variant
// is the fallback for deeper trees
// LDS pointers carry their address space in the type: a pointer that may point at LDS or at
// private memory (the hybrid stack's two halves) would otherwise be merged into one generic
// pointer, and every access would become a flat instruction instead of a ds_ one
typedef __attribute__((address_space(3))) int lds_int;
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) const u32x4 lds_u32x4;
typedef __attribute__((address_space(3))) const u32x2 lds_u32x2;
typedef __attribute__((address_space(3))) const int lds_cint;
struct StackLds {
    lds_int* p;
    RRH_D void put(int i, int v) { p[i * 64] = v; }
    RRH_D int get(int i) const { return p[i * 64]; }
};
struct StackPriv {
    int s[RRH_STACK];
    RRH_D void put(int i, int v) { s[i] = v; }
    RRH_D int get(int i) const { return s[i]; }
};
// the top N entries in LDS, the rest (rarely reached) in private memory
template <int N>
struct StackHybridN {
    lds_int* p;
    int ov[RRH_STACK - N];
    RRH_D void put(int i, int v) {
        if (i < N) p[i * 64] = v;
        else ov[i - N] = v;
    }
    RRH_D int get(int i) const { return (i < N) ? p[i * 64] : ov[i - N]; }
};
using StackHybrid = StackHybridN<RRH_HYB_LDS>;

// stack kinds selectable per context (RRH_OPT_TRAVERSAL_STACK)
template <int K> struct StackSel;
template <> struct StackSel<RRH_STACK_PRIVATE> { using T = StackPriv; static constexpr int kLds = 1; };
template <> struct StackSel<RRH_STACK_LDS32> { using T = StackLds; static constexpr int kLds = 32; };
template <> struct StackSel<RRH_STACK_HYBRID16> { using T = StackHybrid; static constexpr int kLds = RRH_HYB_LDS; };
template <> struct StackSel<RRH_STACK_HYBRID32> { using T = StackHybridN<32>; static constexpr int kLds = 32; };
struct StackNone {};  // stackless traversal (trace_closest4sl_s): no stack memory
template <> struct StackSel<RRH_STACK_NONE> { using T = StackNone; static constexpr int kLds = 1; };

// per-lane LDS base inside a block array of (blockDim/64) x entries x 64 ints
RRH_D lds_int* lds_stack_base(int* smem, int entries) {
    return (lds_int*)(smem + (threadIdx.x >> 6) * (entries * 64) + (threadIdx.x & 63));
}
template <int K>
RRH_D typename StackSel<K>::T make_stack(int* smem) {
    typename StackSel<K>::T st;
    if constexpr (K != RRH_STACK_PRIVATE && K != RRH_STACK_NONE) st.p = lds_stack_base(smem, StackSel<K>::kLds);
    return st;
}

// closest hit over the scene BVH; the result is independent of traversal order
// (ties on t resolve to the smaller flattened id).  COUNT = true is the
// instrumented variant that counts node and triangle fetches.
template <bool COUNT, class Stk>
RRH_D Hit trace_closest2_s(const DevScene& S, f3 o, f3 d, float tmin, float tmax, Stk& stack,
                           uint32_t* n_nodes, uint32_t* n_tris) {
    Hit h; h.t = tmax; h.prim = ~0u; h.b1 = 0.0f; h.b2 = 0.0f;
    WRay r = wray_setup(o, d);
    int sp = 0;
    int node = 0;
    while (true) {
        while (node >= 0) {
            const float4* n = S.nodes + (size_t)node * 4;
            if constexpr (COUNT) (*n_nodes)++;
            float4 a = n[0], b = n[1], c = n[2];
            int4 lk = *reinterpret_cast<const int4*>(n + 3);
            float tl, tr;
            bool hl = box_hit(a.x, a.y, a.z, a.w, c.x, c.y, r, tmin, h.t, tl);
            bool hr = box_hit(b.x, b.y, b.z, b.w, c.z, c.w, r, tmin, h.t, tr);
            if (hl && hr) {
                int nearc = lk.x, farc = lk.y;
                if (tr < tl) { nearc = lk.y; farc = lk.x; }
                stack.put(sp++, farc);
                node = nearc;
            } else if (hl) {
                node = lk.x;
            } else if (hr) {
                node = lk.y;
            } else {
                if (sp == 0) return h;
                node = stack.get(--sp);
            }
        }
        // leaf
        uint32_t first = RRH_LEAF_FIRST(node), cnt = RRH_LEAF_COUNT(node);
        if constexpr (COUNT) (*n_tris) += cnt;
        for (uint32_t i = 0; i < cnt; ++i) {
            const float4* tv = S.tris + (size_t)(first + i) * 3;
            float4 p0 = tv[0], p1 = tv[1], p2 = tv[2];
            float tt, b1, b2;
            uint32_t id = __float_as_uint(p0.w);
            if (wt_hit(r, xyz(p0), xyz(p1), xyz(p2), tmin, h.t, tt, b1, b2) && better(tt, id, h.t, h.prim)) {
                h.t = tt; h.prim = id; h.b1 = b1; h.b2 = b2;
            }
        }
        if (sp == 0) return h;
        node = stack.get(--sp);
    }
}

template <class Stk>
RRH_D bool trace_occluded2_s(const DevScene& S, f3 o, f3 d, float tmin, float tmax, Stk& stack) {
    if (!(tmax >= tmin)) return false;
    WRay r = wray_setup(o, d);
    int sp = 0;
    int node = 0;
    while (true) {
        while (node >= 0) {
            const float4* n = S.nodes + (size_t)node * 4;
            float4 a = n[0], b = n[1], c = n[2];
            int4 lk = *reinterpret_cast<const int4*>(n + 3);
            float tl, tr;
            bool hl = box_hit(a.x, a.y, a.z, a.w, c.x, c.y, r, tmin, tmax, tl);
            bool hr = box_hit(b.x, b.y, b.z, b.w, c.z, c.w, r, tmin, tmax, tr);
            if (hl && hr) {
                stack.put(sp++, lk.y);
                node = lk.x;
            } else if (hl) {
                node = lk.x;
            } else if (hr) {
                node = lk.y;
            } else {
                if (sp == 0) return false;
                node = stack.get(--sp);
            }
        }
        uint32_t first = RRH_LEAF_FIRST(node), cnt = RRH_LEAF_COUNT(node);
        for (uint32_t i = 0; i < cnt; ++i) {
            const float4* tv = S.tris + (size_t)(first + i) * 3;
            float4 p0 = tv[0], p1 = tv[1], p2 = tv[2];
            float tt, b1, b2;
            if (wt_hit(r, xyz(p0), xyz(p1), xyz(p2), tmin, tmax, tt, b1, b2)) return true;
        }
        if (sp == 0) return false;
        node = stack.get(--sp);
    }
}

// ---- speculative binary traversal (Aila and Laine 2009, "Understanding the efficiency of ray
// traversal on GPUs", section 3.3): a lane that reaches a leaf parks it and keeps descending
// inner nodes until every lane of the wave holds a leaf, then the parked leaves are tested
// together, so fewer lanes idle in either loop.  Same closest-hit rule, same answer.
constexpr int kTravDone = 0x7fffffff;  // popped from an empty stack: traversal finished
template <bool COUNT, class Stk>
RRH_D Hit trace_closest2spec_s(const DevScene& S, f3 o, f3 d, float tmin, float tmax, Stk& stack,
                               uint32_t* n_nodes, uint32_t* n_tris) {
    Hit h; h.t = tmax; h.prim = ~0u; h.b1 = 0.0f; h.b2 = 0.0f;
    WRay r = wray_setup(o, d);
    int sp = 0, node = 0, leaf = 0;  // leaf >= 0: none parked
    while (node != kTravDone) {
        while ((uint32_t)node < (uint32_t)kTravDone) {  // an inner node (leaf links are negative)
            const float4* n = S.nodes + (size_t)node * 4;
            if constexpr (COUNT) (*n_nodes)++;
            float4 a = n[0], b = n[1], c = n[2];
            int4 lk = *reinterpret_cast<const int4*>(n + 3);
            float tl, tr;
            bool hl = box_hit(a.x, a.y, a.z, a.w, c.x, c.y, r, tmin, h.t, tl);
            bool hr = box_hit(b.x, b.y, b.z, b.w, c.z, c.w, r, tmin, h.t, tr);
            if (hl && hr) {
                int nearc = lk.x, farc = lk.y;
                if (tr < tl) { nearc = lk.y; farc = lk.x; }
                stack.put(sp++, farc);
                node = nearc;
            } else if (hl) {
                node = lk.x;
            } else if (hr) {
                node = lk.y;
            } else {
                node = (sp == 0) ? kTravDone : stack.get(--sp);
            }
            if (node < 0 && leaf >= 0) {
                leaf = node;
                node = (sp == 0) ? kTravDone : stack.get(--sp);
            }
            if (!__any(leaf >= 0)) break;
        }
        while (leaf < 0) {
            const uint32_t first = RRH_LEAF_FIRST(leaf), cnt = RRH_LEAF_COUNT(leaf);
            if constexpr (COUNT) (*n_tris) += cnt;
            for (uint32_t i = 0; i < cnt; ++i) {
                const float4* tv = S.tris + (size_t)(first + i) * 3;
                float4 p0 = tv[0], p1 = tv[1], p2 = tv[2];
                float tt, b1, b2;
                uint32_t id = __float_as_uint(p0.w);
                if (wt_hit(r, xyz(p0), xyz(p1), xyz(p2), tmin, h.t, tt, b1, b2) && better(tt, id, h.t, h.prim)) {
                    h.t = tt; h.prim = id; h.b1 = b1; h.b2 = b2;
                }
            }
            leaf = node;
            if (node < 0) node = (sp == 0) ? kTravDone : stack.get(--sp);
        }
    }
    return h;
}
template <class Stk>
RRH_D bool trace_occluded2spec_s(const DevScene& S, f3 o, f3 d, float tmin, float tmax, Stk& stack) {
    if (!(tmax >= tmin)) return false;
    WRay r = wray_setup(o, d);
    int sp = 0, node = 0, leaf = 0;
    while (node != kTravDone) {
        while ((uint32_t)node < (uint32_t)kTravDone) {
            const float4* n = S.nodes + (size_t)node * 4;
            float4 a = n[0], b = n[1], c = n[2];
            int4 lk = *reinterpret_cast<const int4*>(n + 3);
            float tl, tr;
            bool hl = box_hit(a.x, a.y, a.z, a.w, c.x, c.y, r, tmin, tmax, tl);
            bool hr = box_hit(b.x, b.y, b.z, b.w, c.z, c.w, r, tmin, tmax, tr);
            if (hl && hr) {
                stack.put(sp++, lk.y);
                node = lk.x;
            } else if (hl) {
                node = lk.x;
            } else if (hr) {
                node = lk.y;
            } else {
                node = (sp == 0) ? kTravDone : stack.get(--sp);
            }
            if (node < 0 && leaf >= 0) {
                leaf = node;
                node = (sp == 0) ? kTravDone : stack.get(--sp);
            }
            if (!__any(leaf >= 0)) break;
        }
        while (leaf < 0) {
            const uint32_t first = RRH_LEAF_FIRST(leaf), cnt = RRH_LEAF_COUNT(leaf);
            for (uint32_t i = 0; i < cnt; ++i) {
                const float4* tv = S.tris + (size_t)(first + i) * 3;
                float4 p0 = tv[0], p1 = tv[1], p2 = tv[2];
                float tt, b1, b2;
                if (wt_hit(r, xyz(p0), xyz(p1), xyz(p2), tmin, tmax, tt, b1, b2)) return true;
            }
            leaf = node;
            if (node < 0) node = (sp == 0) ? kTravDone : stack.get(--sp);
        }
    }
    return false;
}

// ---- stepwise binary traversal for persistent waves ------------------------------------------
// One call runs the inner nodes down to the next leaf and tests that leaf (the while-while
// body); it returns true once the ray is finished.  The state lives in registers between
// calls, so a wave can hand a finished lane a new ray while the others keep going.
struct TravState {
    WRay r;
    float tmin, t;      // t: closest hit so far (closest) / the ray's extent (occluded)
    uint32_t prim;
    float b1, b2;
    int node, sp;
};
RRH_D void trav_begin(TravState& T, f3 o, f3 d, float tmin, float tmax) {
    T.r = wray_setup(o, d);
    T.tmin = tmin; T.t = tmax; T.prim = ~0u; T.b1 = 0.0f; T.b2 = 0.0f; T.node = 0; T.sp = 0;
}
template <bool COUNT, class Stk>
RRH_D bool closest_step(const DevScene& S, TravState& T, Stk& stack, uint32_t& n_nodes, uint32_t& n_tris) {
    while (T.node >= 0) {
        const float4* n = S.nodes + (size_t)T.node * 4;
        if constexpr (COUNT) n_nodes++;
        float4 a = n[0], b = n[1], c = n[2];
        int4 lk = *reinterpret_cast<const int4*>(n + 3);
        float tl, tr;
        bool hl = box_hit(a.x, a.y, a.z, a.w, c.x, c.y, T.r, T.tmin, T.t, tl);
        bool hr = box_hit(b.x, b.y, b.z, b.w, c.z, c.w, T.r, T.tmin, T.t, tr);
        if (hl && hr) {
            int nearc = lk.x, farc = lk.y;
            if (tr < tl) { nearc = lk.y; farc = lk.x; }
            stack.put(T.sp++, farc);
            T.node = nearc;
        } else if (hl) {
            T.node = lk.x;
        } else if (hr) {
            T.node = lk.y;
        } else {
            if (T.sp == 0) return true;
            T.node = stack.get(--T.sp);
        }
    }
    const uint32_t first = RRH_LEAF_FIRST(T.node), cnt = RRH_LEAF_COUNT(T.node);
    if constexpr (COUNT) n_tris += cnt;
    for (uint32_t i = 0; i < cnt; ++i) {
        const float4* tv = S.tris + (size_t)(first + i) * 3;
        float4 p0 = tv[0], p1 = tv[1], p2 = tv[2];
        float tt, b1, b2;
        uint32_t id = __float_as_uint(p0.w);
        if (wt_hit(T.r, xyz(p0), xyz(p1), xyz(p2), T.tmin, T.t, tt, b1, b2) && better(tt, id, T.t, T.prim)) {
            T.t = tt; T.prim = id; T.b1 = b1; T.b2 = b2;
        }
    }
    if (T.sp == 0) return true;
    T.node = stack.get(--T.sp);
    return false;
}
// any hit: returns true once finished; T.prim = 0 marks "occluded"
template <class Stk>
RRH_D bool occluded_step(const DevScene& S, TravState& T, Stk& stack) {
    while (T.node >= 0) {
        const float4* n = S.nodes + (size_t)T.node * 4;
        float4 a = n[0], b = n[1], c = n[2];
        int4 lk = *reinterpret_cast<const int4*>(n + 3);
        float tl, tr;
        bool hl = box_hit(a.x, a.y, a.z, a.w, c.x, c.y, T.r, T.tmin, T.t, tl);
        bool hr = box_hit(b.x, b.y, b.z, b.w, c.z, c.w, T.r, T.tmin, T.t, tr);
        if (hl && hr) {
            stack.put(T.sp++, lk.y);
            T.node = lk.x;
        } else if (hl) {
            T.node = lk.x;
        } else if (hr) {
            T.node = lk.y;
        } else {
            if (T.sp == 0) return true;
            T.node = stack.get(--T.sp);
        }
    }
    const uint32_t first = RRH_LEAF_FIRST(T.node), cnt = RRH_LEAF_COUNT(T.node);
    for (uint32_t i = 0; i < cnt; ++i) {
        const float4* tv = S.tris + (size_t)(first + i) * 3;
        float4 p0 = tv[0], p1 = tv[1], p2 = tv[2];
        float tt, b1, b2;
        if (wt_hit(T.r, xyz(p0), xyz(p1), xyz(p2), T.tmin, T.t, tt, b1, b2)) { T.prim = 0u; return true; }
    }
    if (T.sp == 0) return true;
    T.node = stack.get(--T.sp);
    return false;
}

// wave-level work distribution for persistent waves: hands the idle lanes the next indices
// of a queue of n items (one atomic per refill); `drained` turns true (wave-uniformly) once
// the queue has no more items for this wave
RRH_D bool wave_refill(uint32_t* counter, uint32_t n, bool idle, bool& drained, uint32_t& my) {
    const unsigned long long m = __ballot(idle ? 1 : 0);
    const uint32_t k = (uint32_t)__popcll(m);
    const uint32_t lane = __lane_id();
    const int leader = __ffsll((long long)m) - 1;
    uint32_t base = 0u;
    if ((int)lane == leader) base = atomicAdd(counter, k);
    base = __shfl(base, leader);
    if (base + k >= n) drained = true;
    if (!idle) return false;
    my = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    return my < n;
}

// ---- 4-wide nodes: one 128 B fetch tests four child boxes -----------------------------------
RRH_D float f4c(const float4& v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w)); }
RRH_D int i4c(const int4& v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w)); }
RRH_D void cswap(float& ka, int& va, float& kb, int& vb) {
    if (kb < ka) {
        float tk = ka; ka = kb; kb = tk;
        int tv = va; va = vb; vb = tv;
    }
}

// a 4-wide node as child boxes (fp32 planes) + links; Q: decoded from the quantized form
// (origin + q * 2^(E-127) per bound, the exact fp32 expression the builder checked)
struct Node4 { float4 xl, xh, yl, yh, zl, zh; int4 lk; };
// q * s is exact (an 8-bit integer times a power of two), so the fused form rounds exactly like
// o + q * s with a separate multiply and add
RRH_D float4 dequant4(float o, uint32_t q, float s) {
    return make_float4(__builtin_fmaf((float)(q & 255u), s, o), __builtin_fmaf((float)((q >> 8) & 255u), s, o),
                       __builtin_fmaf((float)((q >> 16) & 255u), s, o), __builtin_fmaf((float)(q >> 24), s, o));
}
// quantized node words: origin (3 floats) and per-axis exponents, then the children's
// 8-bit bounds, then the four links
RRH_D void decode_q(const uint4& a, const uint4& b, const uint4& c, const uint2& d, Node4& N) {
    const float sx = __uint_as_float((a.w & 255u) << 23), sy = __uint_as_float(((a.w >> 8) & 255u) << 23),
                sz = __uint_as_float(((a.w >> 16) & 255u) << 23);
    const float ox = __uint_as_float(a.x), oy = __uint_as_float(a.y), oz = __uint_as_float(a.z);
    N.xl = dequant4(ox, b.x, sx); N.xh = dequant4(ox, b.y, sx);
    N.yl = dequant4(oy, b.z, sy); N.yh = dequant4(oy, b.w, sy);
    N.zl = dequant4(oz, c.x, sz); N.zh = dequant4(oz, c.y, sz);
    N.lk = make_int4((int)c.z, (int)c.w, (int)d.x, (int)d.y);
}
template <bool Q>
RRH_D void load_node4(const DevScene& S, int node, Node4& N) {
    if constexpr (Q) {
        const uint4* n = S.nodesq + (size_t)node * 4;
        decode_q(n[0], n[1], n[2], *reinterpret_cast<const uint2*>(n + 3), N);
    } else {
        const float4* n = S.nodes4 + (size_t)node * 8;
        N.xl = n[0]; N.xh = n[1]; N.yl = n[2]; N.yh = n[3]; N.zl = n[4]; N.zh = n[5];
        N.lk = *reinterpret_cast<const int4*>(n + 6);
    }
}

// Top of the tree in LDS: a workgroup copies the first n quantized nodes into LDS and reads
// them from there.  The nodes are numbered breadth-first at the top of the tree, so those n
// are the root and the levels below it.
struct TopNone {
    static constexpr bool kOn = false;
};
struct TopLds {
    static constexpr bool kOn = true;
    lds_u32x4* p;  // 4 x 16 B per node
    uint32_t n;
};
template <bool Q, class Top>
RRH_D void load_node4(const DevScene& S, const Top& top, int node, Node4& N) {
    if constexpr (Q && Top::kOn) {
        if ((uint32_t)node < top.n) {
            lds_u32x4* n = top.p + (size_t)node * 4;
            const u32x4 a = n[0], b = n[1], c = n[2];
            const u32x2 d = *(lds_u32x2*)(n + 3);
            decode_q(make_uint4(a.x, a.y, a.z, a.w), make_uint4(b.x, b.y, b.z, b.w), make_uint4(c.x, c.y, c.z, c.w),
                     make_uint2(d.x, d.y), N);
            return;
        }
    }
    load_node4<Q>(S, node, N);
}
// one child link of a 4-wide node
template <bool Q, class Top>
RRH_D int node_link(const DevScene& S, const Top& top, int node, uint32_t slot) {
    if constexpr (Q) {
        if constexpr (Top::kOn) {
            if ((uint32_t)node < top.n) return ((lds_cint*)top.p)[(size_t)node * 16 + 10 + slot];
        }
        return reinterpret_cast<const int*>(S.nodesq)[(size_t)node * 16 + 10 + slot];
    } else {
        return reinterpret_cast<const int*>(S.nodes4)[(size_t)node * 32 + 24 + slot];
    }
}
// stages the top n quantized nodes in `lds` (every thread of the block calls it)
template <bool TOP>
RRH_D auto stage_top(const DevScene& S, uint4* lds, uint32_t n) {
    if constexpr (TOP) {
        for (uint32_t i = threadIdx.x; i < n * 4; i += blockDim.x) lds[i] = S.nodesq[i];
        __syncthreads();
        return TopLds{(lds_u32x4*)lds, n};
    } else {
        return TopNone{};
    }
}

// closest hit over the 4-wide nodes: children visited nearest first (the rest pushed far
// to near); the same leaf tests as the binary traversal, so the same answer
// CULL: only triangles facing the ray count (geometric normal against the ray direction),
// the front faces of the reference's G-buffer raster pipeline
// first_hit: return at the first triangle accepted (an occlusion query through the same code)
template <bool COUNT, bool Q, bool CULL = false, class Stk>
RRH_D Hit trace_closest4_s(const DevScene& S, f3 o, f3 d, float tmin, float tmax, Stk& stack,
                           uint32_t* n_nodes, uint32_t* n_tris, bool first_hit = false) {
    Hit h; h.t = tmax; h.prim = ~0u; h.b1 = 0.0f; h.b2 = 0.0f;
    WRay r = wray_setup(o, d);
    int sp = 0;
    int node = 0;
    while (true) {
        while (node >= 0) {
            if constexpr (COUNT) (*n_nodes)++;
            Node4 N;
            load_node4<Q>(S, node, N);
            const float4 xl = N.xl, xh = N.xh, yl = N.yl, yh = N.yh, zl = N.zl, zh = N.zh;
            const int4 lk = N.lk;
            float k[4];
            int c[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float tn;
                const int l = i4c(lk, i);
                const bool hit = box_hit(f4c(xl, i), f4c(xh, i), f4c(yl, i), f4c(yh, i), f4c(zl, i),
                                                     f4c(zh, i), r, tmin, h.t, tn) & (l != 0);
                k[i] = hit ? tn : INFINITY;
                c[i] = l;
            }
            // sorting network on the entry distances (misses sort last)
            cswap(k[0], c[0], k[1], c[1]);
            cswap(k[2], c[2], k[3], c[3]);
            cswap(k[0], c[0], k[2], c[2]);
            cswap(k[1], c[1], k[3], c[3]);
            cswap(k[1], c[1], k[2], c[2]);
            if (!(k[0] < INFINITY)) {
                if (sp == 0) return h;
                node = stack.get(--sp);
                continue;
            }
            if (k[3] < INFINITY) stack.put(sp++, c[3]);
            if (k[2] < INFINITY) stack.put(sp++, c[2]);
            if (k[1] < INFINITY) stack.put(sp++, c[1]);
            node = c[0];
        }
        uint32_t first = RRH_LEAF_FIRST(node), cnt = RRH_LEAF_COUNT(node);
        if constexpr (COUNT) (*n_tris) += cnt;
        for (uint32_t i = 0; i < cnt; ++i) {
            const float4* tv = S.tris + (size_t)(first + i) * 3;
            float4 p0 = tv[0], p1 = tv[1], p2 = tv[2];
            float tt, b1, b2;
            uint32_t id = __float_as_uint(p0.w);
            if (wt_hit(r, xyz(p0), xyz(p1), xyz(p2), tmin, h.t, tt, b1, b2) && better(tt, id, h.t, h.prim)) {
                if constexpr (CULL) {
                    const f3 a = xyz(p0);
                    if (!(dot(cross(sub(xyz(p1), a), sub(xyz(p2), a)), d) < 0.0f)) continue;
                }
                h.t = tt; h.prim = id; h.b1 = b1; h.b2 = b2;
                if (first_hit) return h;
            }
        }
        if (sp == 0) return h;
        node = stack.get(--sp);
    }
}

// CULL: only triangles facing the ray occlude (as in trace_closest4_s)
template <bool Q, bool CULL = false, class Stk>
RRH_D bool trace_occluded4_s(const DevScene& S, f3 o, f3 d, float tmin, float tmax, Stk& stack) {
    if (!(tmax >= tmin)) return false;
    WRay r = wray_setup(o, d);
    int sp = 0;
    int node = 0;
    while (true) {
        while (node >= 0) {
            Node4 N;
            load_node4<Q>(S, node, N);
            const float4 xl = N.xl, xh = N.xh, yl = N.yl, yh = N.yh, zl = N.zl, zh = N.zh;
            const int4 lk = N.lk;
            int next = 0;
            bool have = false;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float tn;
                const int l = i4c(lk, i);
                if (box_hit(f4c(xl, i), f4c(xh, i), f4c(yl, i), f4c(yh, i), f4c(zl, i), f4c(zh, i), r,
                                        tmin, tmax, tn) & (l != 0)) {
                    if (have) stack.put(sp++, next);
                    next = l;
                    have = true;
                }
            }
            if (!have) {
                if (sp == 0) return false;
                node = stack.get(--sp);
                continue;
            }
            node = next;
        }
        uint32_t first = RRH_LEAF_FIRST(node), cnt = RRH_LEAF_COUNT(node);
        for (uint32_t i = 0; i < cnt; ++i) {
            const float4* tv = S.tris + (size_t)(first + i) * 3;
            float4 p0 = tv[0], p1 = tv[1], p2 = tv[2];
            float tt, b1, b2;
            if (wt_hit(r, xyz(p0), xyz(p1), xyz(p2), tmin, tmax, tt, b1, b2)) {
                if constexpr (CULL) {
                    const f3 a = xyz(p0);
                    if (!(dot(cross(sub(xyz(p1), a), sub(xyz(p2), a)), d) < 0.0f)) continue;
                }
                return true;
            }
        }
        if (sp == 0) return false;
        node = stack.get(--sp);
    }
}

// speculative (leaf-parking) traversal over the 4-wide nodes; CULL and first_hit as in
// trace_closest4_s
template <bool COUNT, bool Q, bool CULL = false, class Stk, class Top = TopNone>
RRH_D Hit trace_closest4spec_s(const DevScene& S, f3 o, f3 d, float tmin, float tmax, Stk& stack,
                               uint32_t* n_nodes, uint32_t* n_tris, bool first_hit = false, const Top& top = Top()) {
    Hit h; h.t = tmax; h.prim = ~0u; h.b1 = 0.0f; h.b2 = 0.0f;
    WRay r = wray_setup(o, d);
    int sp = 0, node = 0, leaf = 0;
    while (node != kTravDone) {
        while ((uint32_t)node < (uint32_t)kTravDone) {
            if constexpr (COUNT) (*n_nodes)++;
            Node4 N;
            load_node4<Q>(S, top, node, N);
            float k[4];
            int c[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float tn;
                const int l = i4c(N.lk, i);
                const bool hit = box_hit(f4c(N.xl, i), f4c(N.xh, i), f4c(N.yl, i), f4c(N.yh, i),
                                                     f4c(N.zl, i), f4c(N.zh, i), r, tmin, h.t, tn) & (l != 0);
                k[i] = hit ? tn : INFINITY;
                c[i] = l;
            }
            cswap(k[0], c[0], k[1], c[1]);
            cswap(k[2], c[2], k[3], c[3]);
            cswap(k[0], c[0], k[2], c[2]);
            cswap(k[1], c[1], k[3], c[3]);
            cswap(k[1], c[1], k[2], c[2]);
            if (!(k[0] < INFINITY)) {
                node = (sp == 0) ? kTravDone : stack.get(--sp);
            } else {
                if (k[3] < INFINITY) stack.put(sp++, c[3]);
                if (k[2] < INFINITY) stack.put(sp++, c[2]);
                if (k[1] < INFINITY) stack.put(sp++, c[1]);
                node = c[0];
            }
            if (node < 0 && leaf >= 0) {
                leaf = node;
                node = (sp == 0) ? kTravDone : stack.get(--sp);
            }
            if (!__any(leaf >= 0)) break;
        }
        while (leaf < 0) {
            const uint32_t first = RRH_LEAF_FIRST(leaf), cnt = RRH_LEAF_COUNT(leaf);
            if constexpr (COUNT) (*n_tris) += cnt;
            for (uint32_t i = 0; i < cnt; ++i) {
                const float4* tv = S.tris + (size_t)(first + i) * 3;
                float4 p0 = tv[0], p1 = tv[1], p2 = tv[2];
                float tt, b1, b2;
                uint32_t id = __float_as_uint(p0.w);
                if (wt_hit(r, xyz(p0), xyz(p1), xyz(p2), tmin, h.t, tt, b1, b2) && better(tt, id, h.t, h.prim)) {
                    if constexpr (CULL) {
                        const f3 a = xyz(p0);
                        if (!(dot(cross(sub(xyz(p1), a), sub(xyz(p2), a)), d) < 0.0f)) continue;
                    }
                    h.t = tt; h.prim = id; h.b1 = b1; h.b2 = b2;
                    if (first_hit) return h;
                }
            }
            leaf = node;
            if (node < 0) node = (sp == 0) ? kTravDone : stack.get(--sp);
        }
    }
    return h;
}
template <bool Q, bool COUNT = false, class Stk, class Top = TopNone>
RRH_D bool trace_occluded4spec_s(const DevScene& S, f3 o, f3 d, float tmin, float tmax, Stk& stack,
                                 const Top& top = Top(), uint32_t* n_nodes = nullptr, uint32_t* n_tris = nullptr) {
    if (!(tmax >= tmin)) return false;
    WRay r = wray_setup(o, d);
    int sp = 0, node = 0, leaf = 0;
    while (node != kTravDone) {
        while ((uint32_t)node < (uint32_t)kTravDone) {
            if constexpr (COUNT) (*n_nodes)++;
            Node4 N;
            load_node4<Q>(S, top, node, N);
            int next = kTravDone;
            bool have = false;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float tn;
                const int l = i4c(N.lk, i);
                if (box_hit(f4c(N.xl, i), f4c(N.xh, i), f4c(N.yl, i), f4c(N.yh, i), f4c(N.zl, i),
                                        f4c(N.zh, i), r, tmin, tmax, tn) & (l != 0)) {
                    if (have) stack.put(sp++, next);
                    next = l;
                    have = true;
                }
            }
            node = have ? next : ((sp == 0) ? kTravDone : stack.get(--sp));
            if (node < 0 && leaf >= 0) {
                leaf = node;
                node = (sp == 0) ? kTravDone : stack.get(--sp);
            }
            if (!__any(leaf >= 0)) break;
        }
        while (leaf < 0) {
            const uint32_t first = RRH_LEAF_FIRST(leaf), cnt = RRH_LEAF_COUNT(leaf);
            if constexpr (COUNT) (*n_tris) += cnt;
            for (uint32_t i = 0; i < cnt; ++i) {
                const float4* tv = S.tris + (size_t)(first + i) * 3;
                float4 p0 = tv[0], p1 = tv[1], p2 = tv[2];
                float tt, b1, b2;
                if (wt_hit(r, xyz(p0), xyz(p1), xyz(p2), tmin, tmax, tt, b1, b2)) return true;
            }
            leaf = node;
            if (node < 0) node = (sp == 0) ? kTravDone : stack.get(--sp);
        }
    }
    return false;
}

// ---- resumable speculative traversal of the quantized 4-wide tree ---------------------------
// One ray from where it stands (node, sp, the stack, the hit so far in t / prim / b1 / b2): outer
// steps (a speculative descent until every lane in it holds a leaf, then the parked leaves'
// tests) until the ray is finished (true) or until, at the start of a step, it has visited
// `budget` nodes in this call (false; node, sp and the stack then hold where to continue, with no
// leaf parked).  A new ray starts at node 0, sp 0, prim ~0u and t its extent.
// SHADOW: the extent t is fixed, hit children are pushed in link order and the first hit ends the
// ray (prim = 0 marks it).  Otherwise the nearest child first and the nearest hit kept (ties to
// the smaller id); with first_hit the first accepted hit ends the ray (an occlusion query through
// the closest-hit order, as trace_closest4spec_s does).  CULL: only triangles facing the ray
// direction d count.  No answer depends on the visiting order or on where a ray is interrupted.
template <bool SHADOW, bool CULL = false, class Stk>
RRH_D bool trace4q_steps(const DevScene& S, const WRay& r, f3 d, float tmin, bool first_hit, float& t, uint32_t& prim,
                         float& b1, float& b2, int& node, int& sp, Stk& st, uint32_t budget) {
    int leaf = 0;
    uint32_t visits = 0;
    while (node != kTravDone) {
        if (visits >= budget) return false;
        while ((uint32_t)node < (uint32_t)kTravDone) {
            ++visits;
            Node4 N;
            load_node4<true>(S, node, N);
            if constexpr (SHADOW) {
                int next = kTravDone;
                bool have = false;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    float tn;
                    const int l = i4c(N.lk, i);
                    if (box_hit(f4c(N.xl, i), f4c(N.xh, i), f4c(N.yl, i), f4c(N.yh, i), f4c(N.zl, i), f4c(N.zh, i), r,
                                tmin, t, tn) & (l != 0)) {
                        if (have) st.put(sp++, next);
                        next = l;
                        have = true;
                    }
                }
                node = have ? next : ((sp == 0) ? kTravDone : st.get(--sp));
            } else {
                float k[4];
                int c[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    float tn;
                    const int l = i4c(N.lk, i);
                    const bool hit = box_hit(f4c(N.xl, i), f4c(N.xh, i), f4c(N.yl, i), f4c(N.yh, i), f4c(N.zl, i),
                                             f4c(N.zh, i), r, tmin, t, tn) & (l != 0);
                    k[i] = hit ? tn : INFINITY;
                    c[i] = l;
                }
                cswap(k[0], c[0], k[1], c[1]);
                cswap(k[2], c[2], k[3], c[3]);
                cswap(k[0], c[0], k[2], c[2]);
                cswap(k[1], c[1], k[3], c[3]);
                cswap(k[1], c[1], k[2], c[2]);
                if (!(k[0] < INFINITY)) {
                    node = (sp == 0) ? kTravDone : st.get(--sp);
                } else {
                    if (k[3] < INFINITY) st.put(sp++, c[3]);
                    if (k[2] < INFINITY) st.put(sp++, c[2]);
                    if (k[1] < INFINITY) st.put(sp++, c[1]);
                    node = c[0];
                }
            }
            if (node < 0 && leaf >= 0) {
                leaf = node;
                node = (sp == 0) ? kTravDone : st.get(--sp);
            }
            if (!__any(leaf >= 0)) break;
        }
        // the parked leaf, and a second one the descent stopped at
        while (leaf < 0) {
            const uint32_t first = RRH_LEAF_FIRST(leaf), nt = RRH_LEAF_COUNT(leaf);
            for (uint32_t i = 0; i < nt; ++i) {
                const float4* tv = S.tris + (size_t)(first + i) * 3;
                const float4 p0 = tv[0], p1 = tv[1], p2 = tv[2];
                float tt, u, v;
                if (!wt_hit(r, xyz(p0), xyz(p1), xyz(p2), tmin, t, tt, u, v)) continue;
                const uint32_t id = __float_as_uint(p0.w);
                if constexpr (!SHADOW) {
                    if (!better(tt, id, t, prim)) continue;
                }
                if constexpr (CULL) {
                    const f3 a = xyz(p0);
                    if (!(dot(cross(sub(xyz(p1), a), sub(xyz(p2), a)), d) < 0.0f)) continue;
                }
                if constexpr (SHADOW) {
                    prim = 0u;
                    node = kTravDone;
                    return true;
                } else {
                    t = tt;
                    prim = id;
                    b1 = u;
                    b2 = v;
                    if (first_hit) {
                        node = kTravDone;
                        return true;
                    }
                }
            }
            leaf = node;
            if (node < 0) node = (sp == 0) ? kTravDone : st.get(--sp);
        }
    }
    return true;
}

// ---------------------------------------------------------------------------
// stackless traversal of the 4-wide nodes
// ---------------------------------------------------------------------------
// In place of a stack of node links, each level of the current path keeps 8 bits in
// registers: the number of that node's hit children still to visit (bits 0-1) and their
// slots in visit order (bits 2-7).  When a level runs out the traversal climbs to the parent
// (DevScene::parent4) and drops the level; a pop reads the next slot's link from the node.
// 32 levels in four 64-bit words: trees up to 31 levels deep (the root's level plus one
// spare).  The leaf parking of the speculative traversal is kept; hits are the same as any
// other traversal's (closest t, ties to the smaller id, conservative box tests).
struct Trail {
    uint64_t w0 = 0, w1 = 0, w2 = 0, w3 = 0;
    RRH_D uint32_t level() const { return (uint32_t)w0 & 255u; }
    RRH_D void set_level(uint32_t v) { w0 = (w0 & ~255ull) | v; }
    RRH_D void push(uint32_t v) {
        w3 = (w3 << 8) | (w2 >> 56);
        w2 = (w2 << 8) | (w1 >> 56);
        w1 = (w1 << 8) | (w0 >> 56);
        w0 = (w0 << 8) | v;
    }
    RRH_D void pop() {
        w0 = (w0 >> 8) | (w1 << 56);
        w1 = (w1 >> 8) | (w2 << 56);
        w2 = (w2 >> 8) | (w3 << 56);
        w3 >>= 8;
    }
};
// next node to visit: the next slot of the deepest level with children left (climbing to it),
// or kTravDone at the end
template <bool Q, class Top>
RRH_D int trail_next(const DevScene& S, const Top& top, int& cur, Trail& tr) {
    while (true) {
        const uint32_t b = tr.level();
        const uint32_t left = b & 3u;
        if (left) {
            tr.set_level(((b >> 4) << 2) | (left - 1u));
            return node_link<Q>(S, top, cur, (b >> 2) & 3u);
        }
        if (cur == 0) return kTravDone;
        cur = (int)S.parent4[cur];
        tr.pop();
    }
}

template <bool COUNT, bool Q, bool CULL = false, class Top = TopNone>
RRH_D Hit trace_closest4sl_s(const DevScene& S, f3 o, f3 d, float tmin, float tmax, uint32_t* n_nodes,
                             uint32_t* n_tris, bool first_hit = false, const Top& top = Top()) {
    Hit h; h.t = tmax; h.prim = ~0u; h.b1 = 0.0f; h.b2 = 0.0f;
    WRay r = wray_setup(o, d);
    Trail tr;
    int cur = 0, node = 0, leaf = 0;
    while (node != kTravDone) {
        while ((uint32_t)node < (uint32_t)kTravDone) {
            if constexpr (COUNT) (*n_nodes)++;
            Node4 N;
            load_node4<Q>(S, top, node, N);
            float k[4];
            int c[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float tn;
                const int l = i4c(N.lk, i);
                const bool hit = box_hit(f4c(N.xl, i), f4c(N.xh, i), f4c(N.yl, i), f4c(N.yh, i),
                                                     f4c(N.zl, i), f4c(N.zh, i), r, tmin, h.t, tn) & (l != 0);
                k[i] = hit ? tn : INFINITY;
                c[i] = i;  // slots, sorted by entry distance with their keys
            }
            cswap(k[0], c[0], k[1], c[1]);
            cswap(k[2], c[2], k[3], c[3]);
            cswap(k[0], c[0], k[2], c[2]);
            cswap(k[1], c[1], k[3], c[3]);
            cswap(k[1], c[1], k[2], c[2]);
            if (!(k[0] < INFINITY)) {
                node = trail_next<Q>(S, top, cur, tr);
            } else {
                const uint32_t left = (k[1] < INFINITY ? 1u : 0u) + (k[2] < INFINITY ? 1u : 0u) + (k[3] < INFINITY ? 1u : 0u);
                tr.push(left | (uint32_t)c[1] << 2 | (uint32_t)c[2] << 4 | (uint32_t)c[3] << 6);
                cur = node;
                node = i4c(N.lk, c[0]);
            }
            if (node < 0 && leaf >= 0) {
                leaf = node;
                node = trail_next<Q>(S, top, cur, tr);
            }
            if (!__any(leaf >= 0)) break;
        }
        while (leaf < 0) {
            const uint32_t first = RRH_LEAF_FIRST(leaf), cnt = RRH_LEAF_COUNT(leaf);
            if constexpr (COUNT) (*n_tris) += cnt;
            for (uint32_t i = 0; i < cnt; ++i) {
                const float4* tv = S.tris + (size_t)(first + i) * 3;
                float4 p0 = tv[0], p1 = tv[1], p2 = tv[2];
                float tt, b1, b2;
                uint32_t id = __float_as_uint(p0.w);
                if (wt_hit(r, xyz(p0), xyz(p1), xyz(p2), tmin, h.t, tt, b1, b2) && better(tt, id, h.t, h.prim)) {
                    if constexpr (CULL) {
                        const f3 a = xyz(p0);
                        if (!(dot(cross(sub(xyz(p1), a), sub(xyz(p2), a)), d) < 0.0f)) continue;
                    }
                    h.t = tt; h.prim = id; h.b1 = b1; h.b2 = b2;
                    if (first_hit) return h;
                }
            }
            leaf = node;
            if (node < 0) node = trail_next<Q>(S, top, cur, tr);
        }
    }
    return h;
}
template <bool Q, bool COUNT = false, class Top = TopNone>
RRH_D bool trace_occluded4sl_s(const DevScene& S, f3 o, f3 d, float tmin, float tmax, const Top& top = Top(),
                               uint32_t* n_nodes = nullptr, uint32_t* n_tris = nullptr) {
    if (!(tmax >= tmin)) return false;
    WRay r = wray_setup(o, d);
    Trail tr;
    int cur = 0, node = 0, leaf = 0;
    while (node != kTravDone) {
        while ((uint32_t)node < (uint32_t)kTravDone) {
            if constexpr (COUNT) (*n_nodes)++;
            Node4 N;
            load_node4<Q>(S, top, node, N);
            int first = -1;
            uint32_t lv = 0, left = 0;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float tn;
                const int l = i4c(N.lk, i);
                if (box_hit(f4c(N.xl, i), f4c(N.xh, i), f4c(N.yl, i), f4c(N.yh, i), f4c(N.zl, i),
                                        f4c(N.zh, i), r, tmin, tmax, tn) & (l != 0)) {
                    if (first < 0) {
                        first = i;
                    } else {
                        lv |= (uint32_t)i << (2 + 2 * left);
                        ++left;
                    }
                }
            }
            if (first < 0) {
                node = trail_next<Q>(S, top, cur, tr);
            } else {
                tr.push(lv | left);
                cur = node;
                node = i4c(N.lk, first);
            }
            if (node < 0 && leaf >= 0) {
                leaf = node;
                node = trail_next<Q>(S, top, cur, tr);
            }
            if (!__any(leaf >= 0)) break;
        }
        while (leaf < 0) {
            const uint32_t first = RRH_LEAF_FIRST(leaf), cnt = RRH_LEAF_COUNT(leaf);
            if constexpr (COUNT) (*n_tris) += cnt;
            for (uint32_t i = 0; i < cnt; ++i) {
                const float4* tv = S.tris + (size_t)(first + i) * 3;
                float4 p0 = tv[0], p1 = tv[1], p2 = tv[2];
                float tt, b1, b2;
                if (wt_hit(r, xyz(p0), xyz(p1), xyz(p2), tmin, tmax, tt, b1, b2)) return true;
            }
            leaf = node;
            if (node < 0) node = trail_next<Q>(S, top, cur, tr);
        }
    }
    return false;
}

// the scene's node format (uniform across a launch) picks the traversal; a StackNone stack
// selects the stackless one (the caller has checked that the tree is 4-wide and shallow
// enough for it)
template <bool COUNT, class Stk, class Top = TopNone>
RRH_D Hit trace_closest_s(const DevScene& S, f3 o, f3 d, float tmin, float tmax, Stk& stack,
                          uint32_t* n_nodes = nullptr, uint32_t* n_tris = nullptr, const Top& top = Top()) {
    if constexpr (std::is_same<Stk, StackNone>::value) {
        if (S.node_fmt == RRH_NODES_WIDE4Q)
            return trace_closest4sl_s<COUNT, true>(S, o, d, tmin, tmax, n_nodes, n_tris, false, top);
        return trace_closest4sl_s<COUNT, false>(S, o, d, tmin, tmax, n_nodes, n_tris);
    } else {
        if (S.node_fmt == RRH_NODES_WIDE4Q) {
            if (S.speculative)
                return trace_closest4spec_s<COUNT, true>(S, o, d, tmin, tmax, stack, n_nodes, n_tris, false, top);
            return trace_closest4_s<COUNT, true>(S, o, d, tmin, tmax, stack, n_nodes, n_tris);
        }
        if (S.node_fmt == RRH_NODES_WIDE4) return trace_closest4_s<COUNT, false>(S, o, d, tmin, tmax, stack, n_nodes, n_tris);
        if (S.speculative) return trace_closest2spec_s<COUNT>(S, o, d, tmin, tmax, stack, n_nodes, n_tris);
        return trace_closest2_s<COUNT>(S, o, d, tmin, tmax, stack, n_nodes, n_tris);
    }
}
// COUNT: node and triangle fetches, counted by the stackless and the speculative quantized
// traversals (the bench's configuration); the other paths leave the counters alone
template <bool COUNT = false, class Stk, class Top = TopNone>
RRH_D bool trace_occluded_s(const DevScene& S, f3 o, f3 d, float tmin, float tmax, Stk& stack, const Top& top = Top(),
                            uint32_t* n_nodes = nullptr, uint32_t* n_tris = nullptr) {
    if constexpr (std::is_same<Stk, StackNone>::value) {
        if (S.node_fmt == RRH_NODES_WIDE4Q) return trace_occluded4sl_s<true, COUNT>(S, o, d, tmin, tmax, top, n_nodes, n_tris);
        return trace_occluded4sl_s<false, COUNT>(S, o, d, tmin, tmax, TopNone(), n_nodes, n_tris);
    } else {
        if (S.node_fmt == RRH_NODES_WIDE4Q) {
            if (S.speculative) return trace_occluded4spec_s<true, COUNT>(S, o, d, tmin, tmax, stack, top, n_nodes, n_tris);
            return trace_occluded4_s<true>(S, o, d, tmin, tmax, stack);
        }
        if (S.node_fmt == RRH_NODES_WIDE4) return trace_occluded4_s<false>(S, o, d, tmin, tmax, stack);
        if (S.speculative) return trace_occluded2spec_s(S, o, d, tmin, tmax, stack);
        return trace_occluded2_s(S, o, d, tmin, tmax, stack);
    }
}

// private-stack wrappers (any depth)
template <bool COUNT = false>
RRH_D Hit trace_closest(const DevScene& S, f3 o, f3 d, float tmin, float tmax, uint32_t* n_nodes = nullptr,
                        uint32_t* n_tris = nullptr) {
    StackPriv st;
    return trace_closest_s<COUNT>(S, o, d, tmin, tmax, st, n_nodes, n_tris);
}
RRH_D bool trace_occluded(const DevScene& S, f3 o, f3 d, float tmin, float tmax) {
    StackPriv st;
    return trace_occluded_s(S, o, d, tmin, tmax, st);
}

// XCD-aware block order: the hardware deals consecutive workgroups round-robin to the
// 8 XCDs (each with its own L2).  Remapping gives each XCD runs of kXcdRun consecutive
// logical workgroups (neighbouring tiles or queue ranges, which share cache lines) and
// rotates the runs over the XCDs, so that every XCD still gets a mix of cheap and costly
// image regions; a last partial round of runs keeps the hardware order.
constexpr uint32_t kXcdRun = 16;
RRH_D uint32_t xcd_block(uint32_t b, uint32_t nb, bool remap) {
    if (!remap) return b;
    const uint32_t full = nb - nb % (8u * kXcdRun);
    if (b >= full) return b;
    const uint32_t x = b & 7u, k = b >> 3;
    return ((k / kXcdRun) * 8u + x) * kXcdRun + k % kXcdRun;
}

// ---------------------------------------------------------------------------
// hit attributes, pathtracer.slang:69-117
// ---------------------------------------------------------------------------
struct HitInfo { f3 pos, normal; uint32_t mat; };

RRH_D f3 xform_point(const float* T, f3 p) {
    return mk(mrow(T, 0, p.x, p.y, p.z, 1.0f), mrow(T, 1, p.x, p.y, p.z, 1.0f), mrow(T, 2, p.x, p.y, p.z, 1.0f));
}
RRH_D f3 xform_dir(const float* T, f3 p) {
    return mk(T[0] * p.x + T[1] * p.y + T[2] * p.z, T[4] * p.x + T[5] * p.y + T[6] * p.z,
              T[8] * p.x + T[9] * p.y + T[10] * p.z);
}
// The vertex indices come from the per-triangle record built at upload (vertex_base plus the
// mesh's indices, and the instance), so the vertices are two dependent loads from the hit
// instead of four (instance, mesh, index buffer, vertices).
RRH_D HitInfo fetch_hit(const DevScene& S, uint32_t prim, float b1, float b2) {
    HitInfo hi;
    const uint4 tv = S.tri_vtx[prim];
    const rrh_instance* in = S.insts + tv.w;
    const rrh_mesh* me = S.meshes + in->mesh_index;
    const rrh_vertex* v0 = S.verts + tv.x;
    const rrh_vertex* v1 = S.verts + tv.y;
    const rrh_vertex* v2 = S.verts + tv.z;
    float b0 = 1.0f - b1 - b2;
    f3 lp = add(add(scale(ld3(v0->pos), b0), scale(ld3(v1->pos), b1)), scale(ld3(v2->pos), b2));
    f3 ln = normalize(add(add(scale(ld3(v0->normal), b0), scale(ld3(v1->normal), b1)), scale(ld3(v2->normal), b2)));
    hi.pos = xform_point(in->transform, lp);
    hi.normal = normalize(xform_dir(in->transform, ln));
    hi.mat = me->material_index;
    return hi;
}

// ---------------------------------------------------------------------------
// sampling + BRDF: Sampling.slang:37-66, BRDF.slang:16-187
// ---------------------------------------------------------------------------
#define K_PI 3.14159265358979323846f
#define K_TWO_PI 6.28318530717958647692f
#define K_INV_PI 0.31830988618379067154f
#define K_EPS 1e-4f

RRH_D f3 cosine_hemisphere(float u0, float u1, float& pdf) {
    float r = sqrtf(u0);
    float phi = K_TWO_PI * u1;
    float s, c; sincos_rr(phi, &s, &c);
    float x = r * c, y = r * s;
    float z = sqrtf(fmaxf(0.0f, 1.0f - u0));
    pdf = z * K_INV_PI;
    return mk(x, y, z);
}
RRH_D void build_onb(f3 n, f3& T, f3& B) {
    f3 up = (fabsf(n.x) < 0.9f) ? mk(1.0f, 0.0f, 0.0f) : mk(0.0f, 1.0f, 0.0f);
    T = normalize(cross(up, n));
    B = cross(n, T);
}
RRH_D f3 to_world(f3 l, f3 T, f3 B, f3 N) { return add(add(scale(T, l.x), scale(B, l.y)), scale(N, l.z)); }

RRH_D f3 fresnel_schlick(float cos_theta, f3 F0) {
    float x = clampf(1.0f - cos_theta, 0.0f, 1.0f);
    float x2 = x * x;
    float p = x2 * x2 * x;
    return mk(F0.x + (1.0f - F0.x) * p, F0.y + (1.0f - F0.y) * p, F0.z + (1.0f - F0.z) * p);
}
RRH_D float distribution_ggx(float NdotH, float alpha) {
    float a2 = alpha * alpha;
    float d = NdotH * NdotH * (a2 - 1.0f) + 1.0f;
    return a2 / fmaxf(K_PI * d * d, K_EPS);
}
RRH_D float geometry_smith_ggx(float NdotV, float NdotL, float alpha) {
    float k = alpha * 0.5f;
    float gv = NdotV / fmaxf(NdotV * (1.0f - k) + k, K_EPS);
    float gl = NdotL / fmaxf(NdotL * (1.0f - k) + k, K_EPS);
    return gv * gl;
}
RRH_D f3 brdf_specular_ggx(f3 n, f3 wo, f3 wi, f3 F0, float roughness) {
    float alpha = roughness * roughness;
    f3 h = normalize(add(wo, wi));
    float NdotH = fmaxf(dot(n, h), 0.0f);
    float NdotV = fmaxf(dot(n, wo), 0.0f);
    float NdotL = fmaxf(dot(n, wi), 0.0f);
    float VdotH = fmaxf(dot(wo, h), 0.0f);
    f3 F = fresnel_schlick(VdotH, F0);
    float D = distribution_ggx(NdotH, alpha);
    float G = geometry_smith_ggx(NdotV, NdotL, alpha);
    float den = fmaxf(4.0f * NdotV * NdotL, K_EPS);
    float DG = D * G;
    return mk(DG * F.x / den, DG * F.y / den, DG * F.z / den);
}
RRH_D f3 F0_of(f3 base, float metallic) {
    return mk(lerpf(0.04f, base.x, metallic), lerpf(0.04f, base.y, metallic), lerpf(0.04f, base.z, metallic));
}
RRH_D f3 brdf_eval(f3 n, f3 wo, f3 wi, f3 base, float metallic, float roughness) {
    float r = clampf(roughness, 0.04f, 1.0f);
    f3 F0 = F0_of(base, metallic);
    f3 kS = fresnel_schlick(fmaxf(dot(n, wo), 0.0f), F0);
    f3 kD = mk((1.0f - kS.x) * (1.0f - metallic), (1.0f - kS.y) * (1.0f - metallic), (1.0f - kS.z) * (1.0f - metallic));
    f3 diffuse = mul(kD, scale(base, K_INV_PI));
    f3 spec = brdf_specular_ggx(n, wo, wi, F0, r);
    return add(diffuse, spec);
}
RRH_D f3 sample_ggx_vndf(f3 wo, float roughness, float u0, float u1) {
    float alpha = roughness * roughness;
    f3 Vh = normalize(mk(alpha * wo.x, alpha * wo.y, wo.z));
    float lensq = Vh.x * Vh.x + Vh.y * Vh.y;
    f3 T1;
    if (lensq > 0.0f) { float il = sqrtf(lensq); T1 = mk(-Vh.y / il, Vh.x / il, 0.0f / il); }
    else T1 = mk(1.0f, 0.0f, 0.0f);
    f3 T2 = cross(Vh, T1);
    float r2 = sqrtf(u0);
    float phi = K_TWO_PI * u1;
    float s, c; sincos_rr(phi, &s, &c);
    float t1 = r2 * c, t2 = r2 * s;
    float sv = 0.5f * (1.0f + Vh.z);
    t2 = (1.0f - sv) * sqrtf(1.0f - t1 * t1) + sv * t2;
    float tz = sqrtf(fmaxf(0.0f, 1.0f - t1 * t1 - t2 * t2));
    f3 Nh = add(add(scale(T1, t1), scale(T2, t2)), scale(Vh, tz));
    return normalize(mk(alpha * Nh.x, alpha * Nh.y, fmaxf(0.0f, Nh.z)));
}
RRH_D float p_spec_of(f3 n, f3 wo, f3 base, float metallic) {
    f3 kS = fresnel_schlick(fmaxf(dot(n, wo), 0.0f), F0_of(base, metallic));
    float p = (kS.x + kS.y + kS.z) / 3.0f;
    return clampf(p, 0.1f, 0.9f);
}
RRH_D f3 brdf_sample(f3 wo, f3 n, f3 base, float metallic, float roughness, float ux, float uy, f3& wi,
                     float& pdf) {
    float r = clampf(roughness, 0.04f, 1.0f);
    float p_spec = p_spec_of(n, wo, base, metallic);
    f3 T, B; build_onb(n, T, B);
    f3 wol = mk(dot(wo, T), dot(wo, B), dot(wo, n));
    f3 wil; float pdf_local;
    if (ux < p_spec) {
        f3 h = sample_ggx_vndf(wol, r, ux / p_spec, uy);
        float d2 = 2.0f * dot(h, neg(wol));
        wil = sub(neg(wol), scale(h, d2));
        if (wil.z <= 0.0f) { wi = mk(0.0f, 1.0f, 0.0f); pdf = 1.0f; return mk(0.0f, 0.0f, 0.0f); }
    } else {
        wil = cosine_hemisphere((ux - p_spec) / (1.0f - p_spec), uy, pdf_local);
    }
    wi = to_world(wil, T, B, n);
    float NdotWI = fmaxf(wil.z, 0.0f);
    f3 h = normalize(add(wol, wil));
    float NdotH = fmaxf(h.z, 1e-6f);
    float VdotH = fmaxf(dot(wol, h), 1e-6f);
    float D = distribution_ggx(NdotH, r * r);
    float pdf_spec = D * NdotH / fmaxf(4.0f * VdotH, 1e-6f);
    float pdf_diff = NdotWI * K_INV_PI;
    float p = p_spec * pdf_spec + (1.0f - p_spec) * pdf_diff;
    pdf = fmaxf(p, 1e-6f);
    return brdf_eval(n, wo, wi, base, metallic, roughness);
}
RRH_D float brdf_pdf(f3 wo, f3 wi, f3 n, float roughness) {
    float r = clampf(roughness, 0.04f, 1.0f);
    f3 h = normalize(add(wo, wi));
    float NdotH = fmaxf(dot(n, h), 1e-6f);
    float VdotH = fmaxf(dot(wo, h), 1e-6f);
    float NdotL = fmaxf(dot(n, wi), 0.0f);
    float D = distribution_ggx(NdotH, r * r);
    float pdf_spec = D * NdotH / fmaxf(4.0f * VdotH, 1e-6f);
    float pdf_diff = NdotL * K_INV_PI;
    return 0.5f * pdf_spec + 0.5f * pdf_diff;
}
RRH_D float brdf_sample_pdf(f3 wo, f3 wi, f3 n, f3 base, float metallic, float roughness) {
    float r = clampf(roughness, 0.04f, 1.0f);
    float p_spec = p_spec_of(n, wo, base, metallic);
    f3 h = normalize(add(wo, wi));
    float NdotH = fmaxf(dot(n, h), 1e-6f);
    float VdotH = fmaxf(dot(wo, h), 1e-6f);
    float D = distribution_ggx(NdotH, r * r);
    float pdf_spec = D * NdotH / fmaxf(4.0f * VdotH, 1e-6f);
    float pdf_diff = fmaxf(dot(n, wi), 0.0f) * K_INV_PI;
    return fmaxf(p_spec * pdf_spec + (1.0f - p_spec) * pdf_diff, 1e-6f);
}

struct Mat { f3 base; float metallic, roughness; uint32_t lambert; f3 emissive; };
RRH_D Mat load_mat(const DevScene& S, uint32_t mi) {
    const rrh_material* m = S.mats + mi;
    Mat r;
    r.base = ld3(m->base_color); r.metallic = m->metallic; r.roughness = m->roughness;
    r.lambert = (m->flags & RRH_MAT_PURE_LAMBERT) ? 1u : 0u;
    r.emissive = ld3(m->emissive);
    return r;
}
RRH_D bool is_emissive(const Mat& m) { return m.emissive.x > 0.0f || m.emissive.y > 0.0f || m.emissive.z > 0.0f; }
RRH_D f3 mat_eval(const Mat& m, f3 n, f3 wo, f3 wi) {
    if (m.lambert) return scale(m.base, K_INV_PI);
    return brdf_eval(n, wo, wi, m.base, m.metallic, m.roughness);
}
RRH_D f3 mat_sample(const Mat& m, f3 wo, f3 n, float ux, float uy, f3& wi, float& pdf) {
    if (m.lambert) {
        f3 T, B; build_onb(n, T, B);
        f3 c = cosine_hemisphere(ux, uy, pdf);
        wi = to_world(c, T, B, n);
        return scale(m.base, K_INV_PI);
    }
    return brdf_sample(wo, n, m.base, m.metallic, m.roughness, ux, uy, wi, pdf);
}
RRH_D float mat_sample_pdf(const Mat& m, f3 wo, f3 wi, f3 n) {
    if (m.lambert) return fmaxf(dot(n, wi), 0.0f) * K_INV_PI;
    return brdf_sample_pdf(wo, wi, n, m.base, m.metallic, m.roughness);
}
RRH_D float mat_ref_pdf(const Mat& m, f3 wo, f3 wi, f3 n) {
    if (m.lambert) return fmaxf(dot(n, wi), 0.0f) * K_INV_PI;
    return brdf_pdf(wo, wi, n, m.roughness);
}

// ---------------------------------------------------------------------------
// triangle lights (new; the reference lights are point / directional only)
// ---------------------------------------------------------------------------
// The area (0.5 |e1 x e2|) and unit normal ((e1 x e2) * (1 / |e1 x e2|)) are evaluated once at
// upload, on the host, with the same fp32 operations in the same order; the records carry
// them, so every use reads the identical values instead of recomputing them.
constexpr uint32_t kTlF4 = 5;
struct TriSample { f3 p, nl; float area; f3 Le; };
RRH_D TriSample tri_light_point(const DevScene& S, uint32_t tl, float u1, float u2) {
    const float4* L = S.tl + (size_t)tl * kTlF4;
    float4 q0 = L[0], q1 = L[1], q2 = L[2], q3 = L[3], q4 = L[4];
    f3 a = xyz(q0), b = xyz(q1), c = xyz(q2);
    float su = sqrtf(u1);
    float b0 = 1.0f - su;
    float b1 = u2 * su;
    float b2 = 1.0f - b0 - b1;
    TriSample r;
    r.p = add(add(scale(a, b0), scale(b, b1)), scale(c, b2));
    r.area = q3.y;
    r.nl = xyz(q4);
    r.Le = mk(q0.w, q1.w, q2.w);
    return r;
}
RRH_D uint32_t quant_uv(float u1, float u2) {
    uint32_t a = (uint32_t)(u1 * 65535.0f + 0.5f), b = (uint32_t)(u2 * 65535.0f + 0.5f);
    return a | (b << 16);
}
RRH_D void dequant_uv(uint32_t q, float& u1, float& u2) {
    u1 = (float)(q & 0xffffu) * (1.0f / 65535.0f);
    u2 = (float)(q >> 16) * (1.0f / 65535.0f);
}
RRH_D float tri_light_pdf(const DevScene& S, uint32_t tl, f3 dir, float t) {
    const float4* L = S.tl + (size_t)tl * kTlF4;
    const float area = L[3].y;
    float cosl = fabsf(dot(xyz(L[4]), dir));
    float N = (float)(S.nl + S.ntl);
    return (t * t) / (cosl * area) / N;
}
// floor(u N) mod N for u in [0, 1): the product rounds to at most N, so one conditional
// subtraction replaces the integer division and gives the same index
RRH_D uint32_t pick_index(float u, uint32_t N) {
    const uint32_t v = (uint32_t)(u * (float)N);
    return v >= N ? v - N : v;
}
RRH_D float lum709(f3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

// ---------------------------------------------------------------------------
// vMF lobe: GuidingGrid.slang:160-189
// ---------------------------------------------------------------------------
RRH_D f3 vmf_sample(f3 mu, float kappa, float u0, float u1, float& pdf) {
    float k = fmaxf(kappa, 1.0e-3f);
    float w = 1.0f + (1.0f / k) * log_rr(u0 + (1.0f - u0) * exp_rr(-2.0f * k));
    w = clampf(w, -1.0f, 1.0f);
    float s = sqrtf(fmaxf(0.0f, 1.0f - w * w));
    float phi = K_TWO_PI * u1;
    float sp, cp; sincos_rr(phi, &sp, &cp);
    f3 local = mk(s * cp, s * sp, w);
    f3 T, B; build_onb(mu, T, B);
    f3 dir = to_world(local, T, B, mu);
    float denom = K_TWO_PI * (1.0f - exp_rr(-2.0f * k));
    pdf = k * exp_rr(k * (w - 1.0f)) / fmaxf(denom, 1.0e-8f);
    return dir;
}
RRH_D float vmf_pdf(f3 mu, float kappa, f3 dir) {
    float k = fmaxf(kappa, 1.0e-3f);
    float w = clampf(dot(mu, dir), -1.0f, 1.0f);
    float denom = K_TWO_PI * (1.0f - exp_rr(-2.0f * k));
    return k * exp_rr(k * (w - 1.0f)) / fmaxf(denom, 1.0e-8f);
}

// wave-aggregated queue push: returns this lane's slot (valid only where pred)
RRH_D uint32_t wave_push(uint32_t* counter, bool pred) {
    unsigned long long mask = __ballot(pred ? 1 : 0);
    if (mask == 0ull) return 0u;
    uint32_t lane = __lane_id();
    uint32_t leader = (uint32_t)(__ffsll((long long)mask) - 1);
    uint32_t base = 0u;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, (int)leader);
    unsigned long long lt = (lane == 0u) ? 0ull : (~0ull >> (64u - lane));
    return base + (uint32_t)__popcll(mask & lt);
}

}  // namespace rrh

// rrh_bvh_gpu.hip -- BVH construction on the device, emitting the node formats of rrh_bvh.cpp
// (binary 64 B, 4-wide fp32 128 B, quantized 4-wide 64 B, leaf-ordered triangles).
//
// Replaces the reference's one-off BLAS/TLAS build (Scene.cpp:449-507,
// AccelStructure.cpp:98-330) by a build that stays on the GPU, so a 1M-triangle scene is
// ready in milliseconds instead of the host builder's second.  Steps, N = triangles:
//  1. triangle boxes and centroid bounds (ordered-integer atomics), 63-bit Morton codes of
//     the centroids, radix sort (rocPRIM) -> N leaf clusters in Morton order;
//  2. PLOC (Meister and Bittner, "Parallel Locally-Ordered Clustering for Bounding Volume
//     Hierarchy Construction", IEEE TVCG 2018): every cluster picks, among the clusters at
//     most r positions away, the one whose union box has the smallest surface area (ties by
//     a symmetric key that orders all candidate pairs totally, so the smallest pair is
//     always mutual and every pass merges); mutual pairs merge into a new node, the
//     survivors are compacted by a prefix sum; repeat until one cluster is left;
//  3. leaf collapse, decided as each node is made (its children are complete then): a
//     subtree of at most max_leaf triangles becomes one leaf when its SAH cost as a leaf is
//     not above its cost as a subtree;
//  4. depth-first layout: each node walks to the root summing its left siblings' inner-node
//     and triangle counts, which gives its binary node index and its leaf's first triangle
//     (the host builder's emission order: parent first, left subtree before right);
//  5. 4-wide collapse level by level: each 4-wide node opens the inner child of largest
//     surface area until it has four children (the host builder's rule); node indices in
//     breadth-first order (children after their parent); child boxes quantized to 8 bits,
//     rounded outward and checked in fp32 exactly as the traversal decodes them;
//  6. validation on the device: every link in range and pointing forward, every leaf
//     position referenced once, every decoded quantized box contains its fp32 box.
// Closest hits do not depend on the tree (smaller t, ties to the smaller triangle id, and
// conservative box tests), so a scene traced through this BVH returns the hits of the host
// builder's BVH bit for bit; only the speed differs.
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "rrh_internal.h"

namespace rrh {
namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int kBlock = 256;
constexpr int kPlocMaxR = 32;

// build node: lo.w = left child (int bits), hi.w = right child; a triangle leaf has left = -1
// and right = the triangle's flattened id
struct __align__(16) PNode {
    float4 lo, hi;
};

__device__ __forceinline__ uint32_t ord_of(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float float_of_ord(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}
// half the surface area (a common factor of every comparison)
__device__ __forceinline__ float half_area(float4 lo, float4 hi) {
    const float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
    return dx * dy + dy * dz + dz * dx;
}
__device__ __forceinline__ uint32_t ld_u(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_u(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t left_of(const PNode& p) { return __float_as_uint(p.lo.w); }
__device__ __forceinline__ uint32_t right_of(const PNode& p) { return __float_as_uint(p.hi.w); }
__device__ __forceinline__ bool is_tri(const PNode& p) { return __float_as_int(p.lo.w) < 0; }
__device__ __forceinline__ int32_t leaf_link_d(uint32_t first, uint32_t count) {
    return (int32_t)~((first << 3) | (count - 1u));
}

// 21 bits spread to every third bit of 63
__device__ __forceinline__ uint64_t spread21(uint32_t x) {
    uint64_t v = x & 0x1fffffu;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

// per-block min / max of three coordinates folded into bound[0..2] / bound[3..5] (ordered
// integers) with six atomics per block; every thread of the block must call it
__device__ void block_bounds(float lo[3], float hi[3], uint32_t* __restrict__ bound) {
    __shared__ float part[kBlock / 64][6];
    for (int off = 32; off > 0; off >>= 1)
        for (int k = 0; k < 3; ++k) {
            lo[k] = fminf(lo[k], __shfl_xor(lo[k], off));
            hi[k] = fmaxf(hi[k], __shfl_xor(hi[k], off));
        }
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 3; ++k) {
            part[threadIdx.x >> 6][k] = lo[k];
            part[threadIdx.x >> 6][3 + k] = hi[k];
        }
    __syncthreads();
    if (threadIdx.x < 6) {
        float v = part[0][threadIdx.x];
        for (int w = 1; w < kBlock / 64; ++w)
            v = threadIdx.x < 3 ? fminf(v, part[w][threadIdx.x]) : fmaxf(v, part[w][threadIdx.x]);
        if (threadIdx.x < 3) atomicMin(&bound[threadIdx.x], ord_of(v));
        else atomicMax(&bound[threadIdx.x], ord_of(v));
    }
}

// ---- 1. boxes, Morton codes -----------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_tri_box(const float* __restrict__ tw, uint32_t n, float4* __restrict__ tbox,
                                                    uint32_t* __restrict__ cbound) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    float cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (i < n) {
        const float* v = tw + (size_t)i * 9;
        float lo[3], hi[3];
        for (int k = 0; k < 3; ++k) {
            lo[k] = fminf(fminf(v[k], v[3 + k]), v[6 + k]);
            hi[k] = fmaxf(fmaxf(v[k], v[3 + k]), v[6 + k]);
            cmin[k] = cmax[k] = 0.5f * (lo[k] + hi[k]);
        }
        tbox[2 * (size_t)i] = make_float4(lo[0], lo[1], lo[2], 0.0f);
        tbox[2 * (size_t)i + 1] = make_float4(hi[0], hi[1], hi[2], 0.0f);
    }
    block_bounds(cmin, cmax, cbound);
}

__global__ void k_init_bounds(uint32_t* cbound) {
    if (threadIdx.x < 3) cbound[threadIdx.x] = ord_of(INFINITY);
    else if (threadIdx.x < 6) cbound[threadIdx.x] = ord_of(-INFINITY);
}

__global__ __launch_bounds__(kBlock) void k_morton(const float4* __restrict__ tbox, uint32_t n,
                                                   const uint32_t* __restrict__ cbound, uint64_t* __restrict__ keys,
                                                   uint32_t* __restrict__ vals) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float4 lo = tbox[2 * (size_t)i], hi = tbox[2 * (size_t)i + 1];
    const float c[3] = {0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
    // one scale for all axes (the largest extent): a flat scene keeps square Morton cells
    float ext = 0.0f;
    for (int k = 0; k < 3; ++k) ext = fmaxf(ext, float_of_ord(cbound[3 + k]) - float_of_ord(cbound[k]));
    uint64_t code = 0;
    for (int k = 0; k < 3; ++k) {
        const float mn = float_of_ord(cbound[k]);
        const float t = (ext > 0.0f) ? (c[k] - mn) / ext : 0.0f;
        const uint32_t q = (t >= 0.0f) ? (uint32_t)fminf(t * 2097152.0f, 2097151.0f) : 0u;  // NaN -> 0
        code |= spread21(q) << (2 - k);
    }
    keys[i] = code;
    vals[i] = i;
}

struct Meta {  // per build node
    uint32_t *parent, *cnt, *cost, *ninner, *mark, *first;
};

// SAH leaf collapse, decided when a node is made (its children are complete by then): a
// subtree of at most max_leaf triangles becomes one leaf when its cost as a leaf (area x
// triangles) is not above its cost as a subtree (area x traversal cost + children's costs)
__device__ __forceinline__ void node_meta(Meta m, uint32_t k, float A, uint32_t L, uint32_t R, bool is_root,
                                          uint32_t max_leaf, float trav) {
    const uint32_t n = m.cnt[L] + m.cnt[R];
    const float ci = A * trav + __uint_as_float(m.cost[L]) + __uint_as_float(m.cost[R]);
    const float cl = A * (float)n;
    const bool leaf = !is_root && n <= max_leaf && cl <= ci;
    m.cnt[k] = n;
    m.cost[k] = __float_as_uint(leaf ? cl : ci);
    m.ninner[k] = leaf ? 0u : 1u + m.ninner[L] + m.ninner[R];
    m.mark[k] = leaf ? 1u : 0u;
}

__global__ __launch_bounds__(kBlock) void k_leaves(const float4* __restrict__ tbox, const uint32_t* __restrict__ order,
                                                   uint32_t n, PNode* __restrict__ nodes, uint32_t* __restrict__ cl, Meta m) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    const uint32_t t = order[s];
    float4 lo = tbox[2 * (size_t)t], hi = tbox[2 * (size_t)t + 1];
    lo.w = __int_as_float(-1);
    hi.w = __uint_as_float(t);
    nodes[s].lo = lo;
    nodes[s].hi = hi;
    cl[s] = s;
    m.cnt[s] = 1;
    m.cost[s] = __float_as_uint(half_area(lo, hi));
    m.ninner[s] = 0;
    m.mark[s] = 1;
}

// ---- 2. PLOC ----------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_ploc_nn(const PNode* __restrict__ nodes, const uint32_t* __restrict__ cl,
                                                    uint32_t n, int r, uint32_t* __restrict__ nn) {
    __shared__ float4 slo[kBlock + 2 * kPlocMaxR], shi[kBlock + 2 * kPlocMaxR];
    const int b0 = (int)(blockIdx.x * kBlock);
    const int start = b0 - r, cnt = kBlock + 2 * r;
    for (int k = threadIdx.x; k < cnt; k += kBlock) {
        const int j = start + k;
        if (j >= 0 && j < (int)n) {
            const uint32_t c = cl[j];
            slo[k] = nodes[c].lo;
            shi[k] = nodes[c].hi;
        }
    }
    __syncthreads();
    const int i = b0 + (int)threadIdx.x;
    if (i >= (int)n) return;
    const float4 alo = slo[i - start], ahi = shi[i - start];
    // pair key (area, |i - j|, parity of the lower position, lower position): symmetric and
    // distinct for distinct pairs, so the smallest pair is mutual; on equal areas (repeated
    // geometry) neighbours at even lower positions pair up, halving a run of equal boxes per pass
    float best = INFINITY;
    int bj = -1, bd = 0, bm = 0;
    const int j0 = max(i - r, 0), j1 = min(i + r, (int)n - 1);
    for (int j = j0; j <= j1; ++j) {
        if (j == i) continue;
        const float4 blo = slo[j - start], bhi = shi[j - start];
        const float dx = fmaxf(ahi.x, bhi.x) - fminf(alo.x, blo.x);
        const float dy = fmaxf(ahi.y, bhi.y) - fminf(alo.y, blo.y);
        const float dz = fmaxf(ahi.z, bhi.z) - fminf(alo.z, blo.z);
        float a = dx * dy + dy * dz + dz * dx;
        if (!(a <= INFINITY)) a = INFINITY;  // NaN boxes rank last
        const int d = abs(j - i), mn = min(i, j);
        const bool better = bj < 0 || a < best ||
                            (a == best && (d < bd || (d == bd && ((mn & 1) < (bm & 1) ||
                                                                  ((mn & 1) == (bm & 1) && mn < bm)))));
        if (better) {
            best = a;
            bj = j;
            bd = d;
            bm = mn;
        }
    }
    nn[i] = (uint32_t)bj;
}

// merge flag (high word) and survive flag (low word), summed by one exclusive scan
__global__ __launch_bounds__(kBlock) void k_ploc_flags(const uint32_t* __restrict__ nn, uint32_t n,
                                                       uint64_t* __restrict__ f) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = nn[i];
    const bool mutual = nn[j] == i;
    const uint64_t merge = (mutual && i < j) ? 1u : 0u, survive = (mutual && i > j) ? 0u : 1u;
    f[i] = merge << 32 | survive;
}

__global__ __launch_bounds__(kBlock) void k_ploc_apply(PNode* __restrict__ nodes, const uint32_t* __restrict__ cl,
                                                       const uint32_t* __restrict__ nn, const uint64_t* __restrict__ sc,
                                                       uint32_t n, uint32_t next_node, uint32_t* __restrict__ cl_out,
                                                       uint32_t* __restrict__ totals, Meta meta, uint32_t max_leaf,
                                                       float trav) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = nn[i];
    const bool mutual = nn[j] == i;
    const bool merge = mutual && i < j, survive = !(mutual && i > j);
    const uint64_t s = sc[i];
    const uint32_t pos = (uint32_t)s, m = (uint32_t)(s >> 32);
    if (merge) {
        const uint32_t a = cl[i], b = cl[j], k = next_node + m;
        const PNode A = nodes[a], B = nodes[b];
        PNode o;
        o.lo = make_float4(fminf(A.lo.x, B.lo.x), fminf(A.lo.y, B.lo.y), fminf(A.lo.z, B.lo.z), __uint_as_float(a));
        o.hi = make_float4(fmaxf(A.hi.x, B.hi.x), fmaxf(A.hi.y, B.hi.y), fmaxf(A.hi.z, B.hi.z), __uint_as_float(b));
        nodes[k] = o;
        cl_out[pos] = k;
        node_meta(meta, k, half_area(o.lo, o.hi), a, b, n == 2, max_leaf, trav);
    } else if (survive) {
        cl_out[pos] = cl[i];
    }
    if (i == n - 1) {
        totals[0] = pos + (survive ? 1u : 0u);
        totals[1] = m + (merge ? 1u : 0u);
    }
}

// ---- 3. parents --------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_parents(const PNode* __restrict__ nodes, uint32_t n_leaf, uint32_t n_node,
                                                    uint32_t* __restrict__ parent) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k < n_leaf || k >= n_node) return;
    parent[left_of(nodes[k])] = k;
    parent[right_of(nodes[k])] = k;
    if (k == n_node - 1) parent[k] = kNone;
}

// ---- 4. depth-first layout: binary nodes and leaf-ordered triangles --------------------------------
// stats: [0] binary depth (max over leaves), [1] leaves
__global__ __launch_bounds__(kBlock) void k_layout(const PNode* __restrict__ nodes, uint32_t n_node, Meta m,
                                                   const float* __restrict__ tw, float4* __restrict__ bin,
                                                   float4* __restrict__ tris, uint32_t* __restrict__ stats) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n_node) return;
    uint32_t pre = 0, first = 0, depth = 0, c = k;
    for (;;) {
        const uint32_t p = m.parent[c];
        if (p == kNone) break;
        if (m.mark[p]) return;  // inside a leaf: written by that leaf
        const uint32_t L = left_of(nodes[p]);
        if (c != L) {
            pre += m.ninner[L];
            first += m.cnt[L];
        }
        pre += 1;
        ++depth;
        c = p;
    }
    m.first[k] = first;
    if (m.mark[k]) {
        // a leaf: its triangles in left-to-right order
        uint32_t stk[8];
        int sp = 0;
        uint32_t cur = k, o = first;
        for (;;) {
            const PNode P = nodes[cur];
            if (is_tri(P)) {
                const uint32_t t = right_of(P);
                const float* v = tw + (size_t)t * 9;
                tris[3 * (size_t)o] = make_float4(v[0], v[1], v[2], __uint_as_float(t));
                tris[3 * (size_t)o + 1] = make_float4(v[3], v[4], v[5], 0.0f);
                tris[3 * (size_t)o + 2] = make_float4(v[6], v[7], v[8], 0.0f);
                ++o;
                if (sp == 0) break;
                cur = stk[--sp];
            } else {
                stk[sp++] = right_of(P);
                cur = left_of(P);
            }
        }
        atomicMax(&stats[0], depth);
        atomicAdd(&stats[1], 1u);
        return;
    }
    const PNode P = nodes[k];
    const uint32_t L = left_of(P), R = right_of(P);
    const PNode A = nodes[L], B = nodes[R];
    const int32_t ll = m.mark[L] ? leaf_link_d(first, m.cnt[L]) : (int32_t)(pre + 1);
    const int32_t rl = m.mark[R] ? leaf_link_d(first + m.cnt[L], m.cnt[R]) : (int32_t)(pre + 1 + m.ninner[L]);
    float4* o = bin + 4 * (size_t)pre;
    o[0] = make_float4(A.lo.x, A.hi.x, A.lo.y, A.hi.y);
    o[1] = make_float4(B.lo.x, B.hi.x, B.lo.y, B.hi.y);
    o[2] = make_float4(A.lo.z, A.hi.z, B.lo.z, B.hi.z);
    o[3] = make_float4(__int_as_float(ll), __int_as_float(rl), 0.0f, 0.0f);
}

// ---- 5. 4-wide collapse ---------------------------------------------------------------------------
// children of a 4-wide node rooted at binary inner node q: open the inner child of largest area
__device__ int w4_children(const PNode* __restrict__ nodes, const uint32_t* __restrict__ mark, uint32_t q,
                           uint32_t ch[4]) {
    const PNode Q = nodes[q];
    ch[0] = left_of(Q);
    ch[1] = right_of(Q);
    int n = 2;
    while (n < 4) {
        int best = -1;
        float best_a = -1.0f;
        for (int i = 0; i < n; ++i)
            if (!mark[ch[i]]) {
                const float a = half_area(nodes[ch[i]].lo, nodes[ch[i]].hi);
                if (a > best_a) {
                    best = i;
                    best_a = a;
                }
            }
        if (best < 0) break;
        const uint32_t c = ch[best];
        for (int j = n; j > best + 1; --j) ch[j] = ch[j - 1];
        ch[best] = left_of(nodes[c]);
        ch[best + 1] = right_of(nodes[c]);
        ++n;
    }
    return n;
}

__global__ __launch_bounds__(kBlock) void k_w4_count(const PNode* __restrict__ nodes, const uint32_t* __restrict__ mark,
                                                     const uint32_t* __restrict__ bfs, uint32_t b, uint32_t n,
                                                     uint32_t* __restrict__ icnt) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    uint32_t ch[4];
    const int nch = w4_children(nodes, mark, bfs[b + i], ch);
    uint32_t c = 0;
    for (int s = 0; s < nch; ++s) c += mark[ch[s]] ? 0u : 1u;
    icnt[i] = c;
}

__global__ __launch_bounds__(kBlock) void k_w4_queue(const PNode* __restrict__ nodes, const uint32_t* __restrict__ mark,
                                                     uint32_t* __restrict__ bfs, uint32_t b, uint32_t n, uint32_t e,
                                                     const uint32_t* __restrict__ icnt, const uint32_t* __restrict__ ioff,
                                                     uint32_t* __restrict__ cbase, uint32_t* __restrict__ total) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    uint32_t ch[4];
    const int nch = w4_children(nodes, mark, bfs[b + i], ch);
    const uint32_t base = e + ioff[i];
    uint32_t o = 0;
    for (int s = 0; s < nch; ++s)
        if (!mark[ch[s]]) bfs[base + o++] = ch[s];
    cbase[b + i] = base;
    if (i == n - 1) *total = ioff[i] + icnt[i];
}

// one axis of a quantized node (the host builder's quantize_axis, restated): the smallest
// power-of-two step 2^(E-127) with 255 steps covering the node, then per child 8-bit bounds
// rounded outward until the fp32 decode origin + q * step contains the child
__device__ void quantize_axis_d(float origin, float top, const float* lo, const float* hi, int nch, uint32_t& E,
                                uint32_t& qlo, uint32_t& qhi) {
    const double ext = (double)top - (double)origin;
    int e = -126;
    if (ext > 0.0) {
        int ex = 0;
        const double mant = frexp(ext / 255.0, &ex);
        e = (mant == 0.5) ? ex - 1 : ex;
        e = max(-126, min(127, e));
    }
    while (e > -126 && ldexp(255.0, e - 1) >= ext) --e;
    while (e < 127 && ldexp(255.0, e) < ext) ++e;
    for (;; ++e) {
        const float step = ldexpf(1.0f, e);
        bool ok = true;
        qlo = qhi = 0;
        for (int i = 0; i < nch && ok; ++i) {
            const double fl = floor(((double)lo[i] - (double)origin) / (double)step);
            const double fh = ceil(((double)hi[i] - (double)origin) / (double)step);
            int a = (int)fmax(0.0, fmin(255.0, fl)), b = (int)fmax(0.0, fmin(255.0, fh));
            float dl = origin + (float)a * step, dh = origin + (float)b * step;
            while (a > 0 && dl > lo[i]) { --a; dl = origin + (float)a * step; }
            while (b < 255 && dh < hi[i]) { ++b; dh = origin + (float)b * step; }
            if (dl > lo[i] || dh < hi[i]) ok = false;
            qlo |= (uint32_t)a << (8 * i);
            qhi |= (uint32_t)b << (8 * i);
        }
        if (ok || e >= 127) break;
    }
    E = (uint32_t)(e + 127);
}

__device__ void write_w4(float4* __restrict__ o4, uint4* __restrict__ oq, const float (*lo)[4], const float (*hi)[4],
                         const int32_t* links, int nch) {
    for (int k = 0; k < 3; ++k) {
        o4[2 * k] = make_float4(lo[k][0], lo[k][1], lo[k][2], lo[k][3]);
        o4[2 * k + 1] = make_float4(hi[k][0], hi[k][1], hi[k][2], hi[k][3]);
    }
    o4[6] = make_float4(__int_as_float(links[0]), __int_as_float(links[1]), __int_as_float(links[2]),
                        __int_as_float(links[3]));
    o4[7] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    uint32_t w[16] = {};
    for (int k = 0; k < 3; ++k) {
        float org = lo[k][0], top = hi[k][0];
        for (int i = 1; i < nch; ++i) {
            org = fminf(org, lo[k][i]);
            top = fmaxf(top, hi[k][i]);
        }
        uint32_t E, ql, qh;
        quantize_axis_d(org, top, lo[k], hi[k], nch, E, ql, qh);
        w[k] = __float_as_uint(org);
        w[3] |= E << (8 * k);
        w[4 + 2 * k] = ql;
        w[5 + 2 * k] = qh;
    }
    for (int i = 0; i < 4; ++i) w[10 + i] = (uint32_t)links[i];
    for (int q = 0; q < 4; ++q) oq[q] = make_uint4(w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]);
}

__global__ __launch_bounds__(kBlock) void k_w4_emit(const PNode* __restrict__ nodes, Meta m,
                                                    const uint32_t* __restrict__ bfs, const uint32_t* __restrict__ cbase,
                                                    uint32_t n4, float4* __restrict__ nodes4, uint4* __restrict__ nodesq) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n4) return;
    uint32_t ch[4];
    const int nch = w4_children(nodes, m.mark, bfs[i], ch);
    int32_t links[4] = {0, 0, 0, 0};
    float lo[3][4] = {}, hi[3][4] = {};
    uint32_t o = 0;
    for (int s = 0; s < nch; ++s) {
        const uint32_t c = ch[s];
        const PNode C = nodes[c];
        lo[0][s] = C.lo.x; lo[1][s] = C.lo.y; lo[2][s] = C.lo.z;
        hi[0][s] = C.hi.x; hi[1][s] = C.hi.y; hi[2][s] = C.hi.z;
        links[s] = m.mark[c] ? leaf_link_d(m.first[c], m.cnt[c]) : (int32_t)(cbase[i] + o++);
    }
    write_w4(nodes4 + 8 * (size_t)i, nodesq + 4 * (size_t)i, lo, hi, links, nch);
}

// traversal stack bound, deepest level first: a visit pushes at most nch - 1 siblings
__global__ __launch_bounds__(kBlock) void k_w4_stack(const float4* __restrict__ nodes4, uint32_t b, uint32_t e,
                                                     uint32_t* __restrict__ stk) {
    const uint32_t i = b + blockIdx.x * kBlock + threadIdx.x;
    if (i >= e) return;
    const float4 L = nodes4[8 * (size_t)i + 6];
    const int32_t l[4] = {__float_as_int(L.x), __float_as_int(L.y), __float_as_int(L.z), __float_as_int(L.w)};
    uint32_t nch = 0, sub = 0;
    for (int s = 0; s < 4; ++s) {
        if (l[s] == 0) continue;
        ++nch;
        if (l[s] > 0) sub = max(sub, stk[l[s]]);
    }
    stk[i] = (nch - 1) + sub;
}

// one triangle: a root whose children are that leaf (binary: twice; 4-wide: one slot)
__global__ void k_single(const float* __restrict__ tw, float4* __restrict__ bin, float4* __restrict__ nodes4,
                         uint4* __restrict__ nodesq, float4* __restrict__ tris) {
    if (threadIdx.x != 0) return;
    float lo[3][4] = {}, hi[3][4] = {};
    for (int k = 0; k < 3; ++k) {
        lo[k][0] = fminf(fminf(tw[k], tw[3 + k]), tw[6 + k]);
        hi[k][0] = fmaxf(fmaxf(tw[k], tw[3 + k]), tw[6 + k]);
    }
    const int32_t link = leaf_link_d(0, 1);
    bin[0] = make_float4(lo[0][0], hi[0][0], lo[1][0], hi[1][0]);
    bin[1] = bin[0];
    bin[2] = make_float4(lo[2][0], hi[2][0], lo[2][0], hi[2][0]);
    bin[3] = make_float4(__int_as_float(link), __int_as_float(link), 0.0f, 0.0f);
    const int32_t links[4] = {link, 0, 0, 0};
    write_w4(nodes4, nodesq, lo, hi, links, 1);
    tris[0] = make_float4(tw[0], tw[1], tw[2], __uint_as_float(0u));
    tris[1] = make_float4(tw[3], tw[4], tw[5], 0.0f);
    tris[2] = make_float4(tw[6], tw[7], tw[8], 0.0f);
}

// ---- 6. validation ----------------------------------------------------------------------------
// err bits: 1 binary link, 2 binary leaf range, 4 binary coverage, 8 4-wide link, 16 4-wide leaf
// range, 32 4-wide coverage, 64 quantized links, 128 quantized containment
__device__ bool leaf_refs(int32_t l, uint32_t ntri, uint32_t* ref) {
    const uint32_t u = ~(uint32_t)l, first = u >> 3, cnt = (u & 7u) + 1u;
    if ((uint64_t)first + cnt > ntri) return false;
    for (uint32_t t = first; t < first + cnt; ++t) atomicAdd(&ref[t], 1u);
    return true;
}

__global__ __launch_bounds__(kBlock) void k_val_bin(const float4* __restrict__ bin, uint32_t nn, uint32_t ntri,
                                                    uint32_t* __restrict__ ref, uint32_t* __restrict__ err) {
    const uint32_t n = blockIdx.x * kBlock + threadIdx.x;
    if (n >= nn) return;
    const float4 d = bin[4 * (size_t)n + 3];
    const int32_t l[2] = {__float_as_int(d.x), __float_as_int(d.y)};
    for (int k = 0; k < 2; ++k) {
        if (l[k] >= 0) {
            if ((uint32_t)l[k] >= nn || (uint32_t)l[k] <= n) atomicOr(err, 1u);
        } else if (!leaf_refs(l[k], ntri, ref)) {
            atomicOr(err, 2u);
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_val_refs(uint32_t* __restrict__ ref, uint32_t ntri, uint32_t* __restrict__ err,
                                                     uint32_t bit) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= ntri) return;
    if (ref[t] != 1u) atomicOr(err, bit);
    ref[t] = 0;
}

__global__ __launch_bounds__(kBlock) void k_val_w4(const float4* __restrict__ nodes4, const uint4* __restrict__ nodesq,
                                                   uint32_t n4, uint32_t ntri, uint32_t* __restrict__ ref,
                                                   uint32_t* __restrict__ err) {
    const uint32_t n = blockIdx.x * kBlock + threadIdx.x;
    if (n >= n4) return;
    const float4* f = nodes4 + 8 * (size_t)n;
    const float4 L = f[6];
    const int32_t l[4] = {__float_as_int(L.x), __float_as_int(L.y), __float_as_int(L.z), __float_as_int(L.w)};
    uint32_t w[16];
    for (int q = 0; q < 4; ++q) {
        const uint4 v = nodesq[4 * (size_t)n + q];
        w[4 * q] = v.x; w[4 * q + 1] = v.y; w[4 * q + 2] = v.z; w[4 * q + 3] = v.w;
    }
    for (int s = 0; s < 4; ++s) {
        if ((int32_t)w[10 + s] != l[s]) atomicOr(err, 64u);
        if (l[s] == 0) continue;
        if (l[s] > 0) {
            if ((uint32_t)l[s] >= n4 || (uint32_t)l[s] <= n) atomicOr(err, 8u);
        } else if (!leaf_refs(l[s], ntri, ref)) {
            atomicOr(err, 16u);
        }
        for (int k = 0; k < 3; ++k) {
            const float org = __uint_as_float(w[k]);
            const float step = __uint_as_float(((w[3] >> (8 * k)) & 255u) << 23);
            const float dl = org + (float)((w[4 + 2 * k] >> (8 * s)) & 255u) * step;
            const float dh = org + (float)((w[5 + 2 * k] >> (8 * s)) & 255u) * step;
            const float flo = (&f[2 * k].x)[s], fhi = (&f[2 * k + 1].x)[s];
            if (!(dl <= flo && dh >= fhi)) atomicOr(err, 128u);
        }
    }
}

inline uint32_t grid(size_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

struct Tmp {
    DevBuf tbox, cbound, keys, keys2, vals, vals2, nodes, cl0, cl1, nn, flags, scan, meta, bfs, cbase, icnt, ioff,
        rp, small, ref;
    ~Tmp() {
        DevBuf* all[] = {&tbox, &cbound, &keys, &keys2, &vals, &vals2, &nodes, &cl0, &cl1, &nn, &flags,
                         &scan, &meta, &bfs, &cbase, &icnt, &ioff, &rp, &small, &ref};
        for (DevBuf* b : all) b->release();
    }
};

#define RRH_TRY(x)                      \
    do {                                \
        int rc_ = (x);                  \
        if (rc_) return rc_;            \
    } while (0)
#define RRH_PRIM(x)                                                                        \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) {                                                            \
            set_error(std::string("BVH build (rocPRIM): ") + hipGetErrorString(e_));        \
            return RRH_E_HIP;                                                              \
        }                                                                                  \
    } while (0)

int read_back(void* host, const void* dev, size_t bytes, hipStream_t st) {
    RRH_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, st));
    RRH_HIP(hipStreamSynchronize(st));
    return RRH_OK;
}

}  // namespace

int build_bvh_device(rrh_ctx* c, rrh_scene* s, const float* tw, uint32_t max_leaf, float trav_cost, uint32_t radius) {
    const uint32_t n = s->ntri;
    hipStream_t st = c->stream;
    if (n == 0 || n >= (1u << 28)) { set_error("BVH build: triangle count out of range"); return RRH_E_ARG; }
    if (max_leaf < 1 || max_leaf > 8 || !(trav_cost > 0.0f) || radius < 1 || radius > (uint32_t)kPlocMaxR) {
        set_error("BVH build: bad parameters");
        return RRH_E_ARG;
    }
    Tmp T;
    s->refit_ready = false;  // the refit's parent links describe the previous tree
    RRH_TRY(s->tris.ensure((size_t)n * 3 * sizeof(float4)));
    if (n == 1) {
        RRH_TRY(s->nodes.ensure(4 * sizeof(float4)));
        RRH_TRY(s->nodes4.ensure(8 * sizeof(float4)));
        RRH_TRY(s->nodesq.ensure(4 * sizeof(uint4)));
        k_single<<<1, 64, 0, st>>>(tw, (float4*)s->nodes.p, (float4*)s->nodes4.p, (uint4*)s->nodesq.p,
                                   (float4*)s->tris.p);
        RRH_LAUNCH_CHECK();
        s->nnodes = 1; s->depth = 1; s->leaves = 1; s->nnodes4 = 1; s->depth4 = 1; s->stack4 = 0;
        RRH_HIP(hipStreamSynchronize(st));
        return tree_links(c, s, 1);
    }
    const uint32_t n_node = 2 * n - 1, root = n_node - 1;

    // 1. boxes, Morton codes, sort
    RRH_TRY(T.tbox.ensure((size_t)n * 2 * sizeof(float4)));
    RRH_TRY(T.cbound.ensure(8 * sizeof(uint32_t)));
    RRH_TRY(T.keys.ensure((size_t)n * 8));
    RRH_TRY(T.keys2.ensure((size_t)n * 8));
    RRH_TRY(T.vals.ensure((size_t)n * 4));
    RRH_TRY(T.vals2.ensure((size_t)n * 4));
    k_init_bounds<<<1, 64, 0, st>>>((uint32_t*)T.cbound.p);
    k_tri_box<<<grid(n), kBlock, 0, st>>>(tw, n, (float4*)T.tbox.p, (uint32_t*)T.cbound.p);
    k_morton<<<grid(n), kBlock, 0, st>>>((const float4*)T.tbox.p, n, (const uint32_t*)T.cbound.p, (uint64_t*)T.keys.p,
                                         (uint32_t*)T.vals.p);
    RRH_LAUNCH_CHECK();
    size_t sort_bytes = 0, scan64_bytes = 0, scan32_bytes = 0;
    RRH_PRIM(rocprim::radix_sort_pairs(nullptr, sort_bytes, (uint64_t*)T.keys.p, (uint64_t*)T.keys2.p,
                                       (uint32_t*)T.vals.p, (uint32_t*)T.vals2.p, n, 0, 63, st));
    RRH_PRIM(rocprim::exclusive_scan(nullptr, scan64_bytes, (const uint64_t*)T.flags.p, (uint64_t*)T.scan.p,
                                     (uint64_t)0, n, rocprim::plus<uint64_t>(), st));
    RRH_PRIM(rocprim::exclusive_scan(nullptr, scan32_bytes, (const uint32_t*)T.icnt.p, (uint32_t*)T.ioff.p,
                                     (uint32_t)0, n, rocprim::plus<uint32_t>(), st));
    size_t rp_bytes = std::max(sort_bytes, std::max(scan64_bytes, scan32_bytes));
    RRH_TRY(T.rp.ensure(rp_bytes));
    RRH_PRIM(rocprim::radix_sort_pairs(T.rp.p, sort_bytes, (uint64_t*)T.keys.p, (uint64_t*)T.keys2.p,
                                       (uint32_t*)T.vals.p, (uint32_t*)T.vals2.p, n, 0, 63, st));
    T.keys.release();  // sorted keys are not needed, only the order
    T.keys2.release();

    // 2. PLOC
    RRH_TRY(T.nodes.ensure((size_t)n_node * sizeof(PNode)));
    RRH_TRY(T.cl0.ensure((size_t)n * 4));
    RRH_TRY(T.cl1.ensure((size_t)n * 4));
    RRH_TRY(T.nn.ensure((size_t)n * 4));
    RRH_TRY(T.flags.ensure((size_t)n * 8));
    RRH_TRY(T.scan.ensure((size_t)n * 8));
    RRH_TRY(T.small.ensure(32 * sizeof(uint32_t)));  // [0..1] totals, [8..9] layout stats, [16] validation bits
    RRH_TRY(T.meta.ensure((size_t)n_node * 6 * sizeof(uint32_t)));
    Meta m;
    uint32_t* mb = (uint32_t*)T.meta.p;
    m.parent = mb;
    m.cnt = mb + (size_t)n_node;
    m.cost = mb + 2 * (size_t)n_node;
    m.ninner = mb + 3 * (size_t)n_node;
    m.mark = mb + 4 * (size_t)n_node;
    m.first = mb + 5 * (size_t)n_node;
    PNode* nodes = (PNode*)T.nodes.p;
    k_leaves<<<grid(n), kBlock, 0, st>>>((const float4*)T.tbox.p, (const uint32_t*)T.vals2.p, n, nodes,
                                         (uint32_t*)T.cl0.p, m);
    RRH_LAUNCH_CHECK();
    uint32_t* cl = (uint32_t*)T.cl0.p;
    uint32_t* cl_next = (uint32_t*)T.cl1.p;
    uint32_t live = n, next_node = n, passes = 0;
    while (live > 1) {
        k_ploc_nn<<<grid(live), kBlock, 0, st>>>(nodes, cl, live, (int)radius, (uint32_t*)T.nn.p);
        k_ploc_flags<<<grid(live), kBlock, 0, st>>>((const uint32_t*)T.nn.p, live, (uint64_t*)T.flags.p);
        RRH_LAUNCH_CHECK();
        size_t bytes = scan64_bytes;
        RRH_PRIM(rocprim::exclusive_scan(T.rp.p, bytes, (const uint64_t*)T.flags.p, (uint64_t*)T.scan.p, (uint64_t)0,
                                         live, rocprim::plus<uint64_t>(), st));
        k_ploc_apply<<<grid(live), kBlock, 0, st>>>(nodes, cl, (const uint32_t*)T.nn.p, (const uint64_t*)T.scan.p, live,
                                                    next_node, cl_next, (uint32_t*)T.small.p, m, max_leaf, trav_cost);
        RRH_LAUNCH_CHECK();
        uint32_t tot[2];
        RRH_TRY(read_back(tot, T.small.p, 8, st));
        if (tot[1] == 0 || tot[0] + tot[1] != live || next_node + tot[1] > n_node) {
            set_error("BVH build: internal error (clustering made no progress)");
            return RRH_E_STATE;
        }
        live = tot[0];
        next_node += tot[1];
        std::swap(cl, cl_next);
        ++passes;
    }
    if (next_node != n_node) { set_error("BVH build: internal error (node count)"); return RRH_E_STATE; }
    T.cl0.release(); T.cl1.release(); T.nn.release(); T.flags.release(); T.scan.release(); T.vals.release();

    // 3. parents (the leaf collapse was decided as each node was made)
    k_parents<<<grid(n_node), kBlock, 0, st>>>(nodes, n, n_node, m.parent);
    RRH_LAUNCH_CHECK();
    uint32_t nbin = 0;
    RRH_TRY(read_back(&nbin, m.ninner + root, 4, st));
    if (nbin == 0 || nbin >= n) { set_error("BVH build: internal error (inner node count)"); return RRH_E_STATE; }

    // 4. binary nodes and triangles, depth first
    RRH_TRY(s->nodes.ensure((size_t)nbin * 4 * sizeof(float4)));
    uint32_t* stats = (uint32_t*)T.small.p + 8;
    RRH_HIP(hipMemsetAsync(stats, 0, 8 * sizeof(uint32_t), st));
    k_layout<<<grid(n_node), kBlock, 0, st>>>(nodes, n_node, m, tw, (float4*)s->nodes.p, (float4*)s->tris.p, stats);
    RRH_LAUNCH_CHECK();

    // 5. 4-wide nodes, level by level
    RRH_TRY(T.bfs.ensure((size_t)nbin * 4));
    RRH_TRY(T.cbase.ensure((size_t)nbin * 4));
    RRH_TRY(T.icnt.ensure((size_t)nbin * 4));
    RRH_TRY(T.ioff.ensure((size_t)nbin * 4));
    RRH_HIP(hipMemcpyAsync(T.bfs.p, &root, 4, hipMemcpyHostToDevice, st));
    std::vector<uint32_t> level{0u, 1u};  // level l holds nodes [level[l], level[l + 1])
    while (level.back() > level[level.size() - 2]) {
        const uint32_t b = level[level.size() - 2], e = level.back(), cnt = e - b;
        k_w4_count<<<grid(cnt), kBlock, 0, st>>>(nodes, m.mark, (const uint32_t*)T.bfs.p, b, cnt, (uint32_t*)T.icnt.p);
        RRH_LAUNCH_CHECK();
        size_t bytes = scan32_bytes;
        RRH_PRIM(rocprim::exclusive_scan(T.rp.p, bytes, (const uint32_t*)T.icnt.p, (uint32_t*)T.ioff.p, (uint32_t)0, cnt,
                                         rocprim::plus<uint32_t>(), st));
        k_w4_queue<<<grid(cnt), kBlock, 0, st>>>(nodes, m.mark, (uint32_t*)T.bfs.p, b, cnt, e,
                                                 (const uint32_t*)T.icnt.p, (const uint32_t*)T.ioff.p,
                                                 (uint32_t*)T.cbase.p, (uint32_t*)T.small.p);
        RRH_LAUNCH_CHECK();
        uint32_t add = 0;
        RRH_TRY(read_back(&add, T.small.p, 4, st));
        if ((uint64_t)e + add > nbin) { set_error("BVH build: internal error (4-wide node count)"); return RRH_E_STATE; }
        level.push_back(e + add);
    }
    const uint32_t n4 = level.back();
    RRH_TRY(s->nodes4.ensure((size_t)n4 * 8 * sizeof(float4)));
    RRH_TRY(s->nodesq.ensure((size_t)n4 * 4 * sizeof(uint4)));
    k_w4_emit<<<grid(n4), kBlock, 0, st>>>(nodes, m, (const uint32_t*)T.bfs.p, (const uint32_t*)T.cbase.p, n4,
                                           (float4*)s->nodes4.p, (uint4*)s->nodesq.p);
    RRH_LAUNCH_CHECK();
    uint32_t* stk = (uint32_t*)T.icnt.p;  // reused: one entry per 4-wide node
    for (size_t l = level.size() - 1; l-- > 0;) {  // deepest first; the last range is empty
        if (level[l + 1] > level[l])
            k_w4_stack<<<grid(level[l + 1] - level[l]), kBlock, 0, st>>>((const float4*)s->nodes4.p, level[l],
                                                                         level[l + 1], stk);
    }
    RRH_LAUNCH_CHECK();

    // 6. validation
    RRH_TRY(T.ref.ensure((size_t)n * 4));
    RRH_HIP(hipMemsetAsync(T.ref.p, 0, (size_t)n * 4, st));
    uint32_t* err = (uint32_t*)T.small.p + 16;
    RRH_HIP(hipMemsetAsync(err, 0, 4, st));
    k_val_bin<<<grid(nbin), kBlock, 0, st>>>((const float4*)s->nodes.p, nbin, n, (uint32_t*)T.ref.p, err);
    k_val_refs<<<grid(n), kBlock, 0, st>>>((uint32_t*)T.ref.p, n, err, 4u);
    k_val_w4<<<grid(n4), kBlock, 0, st>>>((const float4*)s->nodes4.p, (const uint4*)s->nodesq.p, n4, n,
                                          (uint32_t*)T.ref.p, err);
    k_val_refs<<<grid(n), kBlock, 0, st>>>((uint32_t*)T.ref.p, n, err, 32u);
    RRH_LAUNCH_CHECK();
    uint32_t res[9];
    RRH_TRY(read_back(res, (const char*)T.small.p + 8 * sizeof(uint32_t), 9 * sizeof(uint32_t), st));
    uint32_t stack4 = 0;
    RRH_TRY(read_back(&stack4, stk, 4, st));
    const uint32_t ebits = res[8];
    if (ebits) {
        set_error("BVH build: internal error (device validation bits " + std::to_string(ebits) + ")");
        return RRH_E_STATE;
    }
    s->nnodes = nbin;
    s->depth = res[0];
    s->leaves = res[1];
    s->nnodes4 = n4;
    s->depth4 = (uint32_t)level.size() - 2;
    s->stack4 = stack4;
    if (s->depth > RRH_STACK - 1 || s->stack4 > RRH_STACK) {
        set_error("BVH build: the device-built tree is deeper than the 64-entry traversal stack (depth " +
                  std::to_string(s->depth) + ", 4-wide bound " + std::to_string(s->stack4) +
                  "); use the host builder");
        return RRH_E_STATE;
    }
    s->build_passes = passes;
    return tree_links(c, s, n4);  // every node is numbered breadth-first
}

// ---- instance updates: flattening, light records, refit ----------------------------------------
namespace {

// world position row r of a row-major instance transform (the host flattening's expression)
__device__ __forceinline__ float mrow_d(const float* M, int r, float x, float y, float z) {
    return M[r * 4 + 0] * x + M[r * 4 + 1] * y + M[r * 4 + 2] * z + M[r * 4 + 3] * 1.0f;
}

__global__ __launch_bounds__(kBlock) void k_flatten(const uint2* __restrict__ tri_ref, const rrh_instance* __restrict__ insts,
                                                    const rrh_mesh* __restrict__ meshes, const uint32_t* __restrict__ idx,
                                                    const rrh_vertex* __restrict__ verts, uint32_t ntri,
                                                    float* __restrict__ tw, uint32_t* __restrict__ bound) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (t < ntri) {
        const uint2 r = tri_ref[t];
        const rrh_instance& I = insts[r.x];
        const rrh_mesh& me = meshes[I.mesh_index];
        const uint32_t base = me.index_byte_offset / 4 + r.y * 3;
        float* w = tw + (size_t)t * 9;
        for (int k = 0; k < 3; ++k) {
            const rrh_vertex& v = verts[me.vertex_base + idx[base + k]];
            for (int a = 0; a < 3; ++a) {
                const float x = mrow_d(I.transform, a, v.pos[0], v.pos[1], v.pos[2]);
                w[k * 3 + a] = x;
                lo[a] = fminf(lo[a], x);
                hi[a] = fmaxf(hi[a], x);
            }
        }
    }
    block_bounds(lo, hi, bound);
}

// triangle-light records from the moved triangles (the upload's fp32 expressions, same order)
__global__ __launch_bounds__(kBlock) void k_tl_update(const float* __restrict__ tw, uint32_t ntl, float4* __restrict__ tl) {
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= ntl) return;
    float4* L = tl + (size_t)j * kTlF4;
    const uint32_t t = __float_as_uint(L[3].x);
    const float* w = tw + (size_t)t * 9;
    L[0] = make_float4(w[0], w[1], w[2], L[0].w);
    L[1] = make_float4(w[3], w[4], w[5], L[1].w);
    L[2] = make_float4(w[6], w[7], w[8], L[2].w);
    const float e1x = w[3] - w[0], e1y = w[4] - w[1], e1z = w[5] - w[2];
    const float e2x = w[6] - w[0], e2y = w[7] - w[1], e2z = w[8] - w[2];
    const float crx = e1y * e2z - e1z * e2y, cry = e1z * e2x - e1x * e2z, crz = e1x * e2y - e1y * e2x;
    const float l2 = sqrtf(crx * crx + cry * cry + crz * crz);
    const float inv = 1.0f / l2;
    L[3].y = 0.5f * l2;
    L[4] = make_float4(crx * inv, cry * inv, crz * inv, 0.0f);
}

// leaf-ordered triangles from the moved world triangles (ids kept)
__global__ __launch_bounds__(kBlock) void k_tris_refresh(const float* __restrict__ tw, uint32_t ntri, float4* __restrict__ tris) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= ntri) return;
    float4* o = tris + 3 * (size_t)k;
    const uint32_t t = __float_as_uint(o[0].w);
    const float* v = tw + (size_t)t * 9;
    o[0] = make_float4(v[0], v[1], v[2], __uint_as_float(t));
    o[1] = make_float4(v[3], v[4], v[5], 0.0f);
    o[2] = make_float4(v[6], v[7], v[8], 0.0f);
}

// box of a leaf link: the union of its triangles' vertices
__device__ void leaf_box(const float4* __restrict__ tris, int32_t link, float lo[3], float hi[3]) {
    const uint32_t u = ~(uint32_t)link, first = u >> 3, cnt = (u & 7u) + 1u;
    for (int k = 0; k < 3; ++k) { lo[k] = INFINITY; hi[k] = -INFINITY; }
    for (uint32_t i = first; i < first + cnt; ++i)
        for (int v = 0; v < 3; ++v) {
            const float4 p = tris[3 * (size_t)i + v];
            lo[0] = fminf(lo[0], p.x); lo[1] = fminf(lo[1], p.y); lo[2] = fminf(lo[2], p.z);
            hi[0] = fmaxf(hi[0], p.x); hi[1] = fmaxf(hi[1], p.y); hi[2] = fmaxf(hi[2], p.z);
        }
}

__device__ __forceinline__ void st_f(float* p, float v) { st_u(reinterpret_cast<uint32_t*>(p), __float_as_uint(v)); }
__device__ __forceinline__ float ld_f(const float* p) { return __uint_as_float(ld_u(reinterpret_cast<const uint32_t*>(p))); }

// parent links: parent[child] = (node << shift) | slot; inner[node] = inner children
__global__ __launch_bounds__(kBlock) void k_bin_parents(const float4* __restrict__ bin, uint32_t nn, uint32_t* __restrict__ parent,
                                                        uint32_t* __restrict__ inner) {
    const uint32_t n = blockIdx.x * kBlock + threadIdx.x;
    if (n >= nn) return;
    const float4 d = bin[4 * (size_t)n + 3];
    const int32_t l[2] = {__float_as_int(d.x), __float_as_int(d.y)};
    uint32_t c = 0;
    for (uint32_t k = 0; k < 2; ++k)
        if (l[k] >= 0) { parent[l[k]] = (n << 1) | k; ++c; }
    inner[n] = c;
    if (n == 0) parent[0] = kNone;
}
__global__ __launch_bounds__(kBlock) void k_w4_parents(const float4* __restrict__ nodes4, uint32_t n4, uint32_t* __restrict__ parent,
                                                       uint32_t* __restrict__ inner) {
    const uint32_t n = blockIdx.x * kBlock + threadIdx.x;
    if (n >= n4) return;
    const float4 L = nodes4[8 * (size_t)n + 6];
    const int32_t l[4] = {__float_as_int(L.x), __float_as_int(L.y), __float_as_int(L.z), __float_as_int(L.w)};
    uint32_t c = 0;
    for (uint32_t s = 0; s < 4; ++s)
        if (l[s] > 0) { parent[l[s]] = (n << 2) | s; ++c; }
    inner[n] = c;
    if (n == 0) parent[0] = kNone;
}

// parent of every 4-wide node (the root's is itself, never read)
__global__ __launch_bounds__(kBlock) void k_w4_parent_index(const float4* __restrict__ nodes4, uint32_t n4,
                                                            uint32_t* __restrict__ parent) {
    const uint32_t n = blockIdx.x * kBlock + threadIdx.x;
    if (n >= n4) return;
    const float4 L = nodes4[8 * (size_t)n + 6];
    const int32_t l[4] = {__float_as_int(L.x), __float_as_int(L.y), __float_as_int(L.z), __float_as_int(L.w)};
    for (uint32_t k = 0; k < 4; ++k)
        if (l[k] > 0 && (uint32_t)l[k] < n4) parent[l[k]] = n;
    if (n == 0) parent[0] = 0;
}

// binary refit, bottom-up: a node whose children are all leaves starts; the last child to
// finish a node completes it (leaf slots from the triangles, inner slots already written)
__global__ __launch_bounds__(kBlock) void k_bin_refit(float4* __restrict__ bin, uint32_t nn, const float4* __restrict__ tris,
                                                      const uint32_t* __restrict__ parent, const uint32_t* __restrict__ inner,
                                                      uint32_t* __restrict__ arrive) {
    const uint32_t n = blockIdx.x * kBlock + threadIdx.x;
    if (n >= nn || inner[n] != 0) return;
    uint32_t cur = n;
    for (;;) {
        float* f = reinterpret_cast<float*>(bin + 4 * (size_t)cur);
        const int32_t l[2] = {__float_as_int(f[12]), __float_as_int(f[13])};
        // slot k: lo.x hi.x lo.y hi.y at f[4k..4k+3], lo.z hi.z at f[8 + 2k], f[9 + 2k]
        float blo[3] = {INFINITY, INFINITY, INFINITY}, bhi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int k = 0; k < 2; ++k) {
            float lo[3], hi[3];
            if (l[k] < 0) {
                leaf_box(tris, l[k], lo, hi);
                st_f(&f[4 * k], lo[0]); st_f(&f[4 * k + 1], hi[0]); st_f(&f[4 * k + 2], lo[1]); st_f(&f[4 * k + 3], hi[1]);
                st_f(&f[8 + 2 * k], lo[2]); st_f(&f[9 + 2 * k], hi[2]);
            } else {
                lo[0] = ld_f(&f[4 * k]); hi[0] = ld_f(&f[4 * k + 1]); lo[1] = ld_f(&f[4 * k + 2]); hi[1] = ld_f(&f[4 * k + 3]);
                lo[2] = ld_f(&f[8 + 2 * k]); hi[2] = ld_f(&f[9 + 2 * k]);
            }
            for (int a = 0; a < 3; ++a) { blo[a] = fminf(blo[a], lo[a]); bhi[a] = fmaxf(bhi[a], hi[a]); }
        }
        const uint32_t p = parent[cur];
        if (p == kNone) return;
        const uint32_t pn = p >> 1, k = p & 1u;
        float* g = reinterpret_cast<float*>(bin + 4 * (size_t)pn);
        st_f(&g[4 * k], blo[0]); st_f(&g[4 * k + 1], bhi[0]); st_f(&g[4 * k + 2], blo[1]); st_f(&g[4 * k + 3], bhi[1]);
        st_f(&g[8 + 2 * k], blo[2]); st_f(&g[9 + 2 * k], bhi[2]);
        __threadfence();
        if (atomicAdd(&arrive[pn], 1u) + 1u < inner[pn]) return;
        __threadfence();
        cur = pn;
    }
}

// 4-wide refit: the same walk; a completed node rewrites its fp32 planes and its quantized copy
__global__ __launch_bounds__(kBlock) void k_w4_refit(float4* __restrict__ nodes4, uint4* __restrict__ nodesq, uint32_t n4,
                                                     const float4* __restrict__ tris, const uint32_t* __restrict__ parent,
                                                     const uint32_t* __restrict__ inner, uint32_t* __restrict__ arrive) {
    const uint32_t n = blockIdx.x * kBlock + threadIdx.x;
    if (n >= n4 || inner[n] != 0) return;
    uint32_t cur = n;
    for (;;) {
        float* f = reinterpret_cast<float*>(nodes4 + 8 * (size_t)cur);
        int32_t links[4];
        for (int s = 0; s < 4; ++s) links[s] = __float_as_int(f[24 + s]);
        float lo[3][4] = {}, hi[3][4] = {};
        float blo[3] = {INFINITY, INFINITY, INFINITY}, bhi[3] = {-INFINITY, -INFINITY, -INFINITY};
        int nch = 0;
        for (int s = 0; s < 4; ++s) {
            if (links[s] == 0) continue;
            nch = s + 1;
            float a[3], b[3];
            if (links[s] < 0) {
                leaf_box(tris, links[s], a, b);
            } else {
                for (int k = 0; k < 3; ++k) { a[k] = ld_f(&f[8 * k + s]); b[k] = ld_f(&f[8 * k + 4 + s]); }
            }
            for (int k = 0; k < 3; ++k) {
                lo[k][s] = a[k];
                hi[k][s] = b[k];
                blo[k] = fminf(blo[k], a[k]);
                bhi[k] = fmaxf(bhi[k], b[k]);
            }
        }
        write_w4(nodes4 + 8 * (size_t)cur, nodesq + 4 * (size_t)cur, lo, hi, links, nch);
        const uint32_t p = parent[cur];
        if (p == kNone) return;
        const uint32_t pn = p >> 2, s = p & 3u;
        float* g = reinterpret_cast<float*>(nodes4 + 8 * (size_t)pn);
        for (int k = 0; k < 3; ++k) { st_f(&g[8 * k + s], blo[k]); st_f(&g[8 * k + 4 + s], bhi[k]); }
        __threadfence();
        if (atomicAdd(&arrive[pn], 1u) + 1u < inner[pn]) return;
        __threadfence();
        cur = pn;
    }
}

}  // namespace

int flatten_device(rrh_ctx* c, rrh_scene* s) {
    hipStream_t st = c->stream;
    const uint32_t n = s->ntri;
    RRH_TRY(s->twd.ensure((size_t)n * 9 * sizeof(float)));
    DevBuf bound;
    RRH_TRY(bound.ensure(8 * sizeof(uint32_t)));
    k_init_bounds<<<1, 64, 0, st>>>((uint32_t*)bound.p);
    k_flatten<<<grid(n), kBlock, 0, st>>>((const uint2*)s->tri_ref.p, (const rrh_instance*)s->insts.p,
                                          (const rrh_mesh*)s->meshes.p, (const uint32_t*)s->idx.p,
                                          (const rrh_vertex*)s->verts.p, n, (float*)s->twd.p, (uint32_t*)bound.p);
    if (s->ntl) k_tl_update<<<grid(s->ntl), kBlock, 0, st>>>((const float*)s->twd.p, s->ntl, (float4*)s->tl.p);
    RRH_LAUNCH_CHECK();
    uint32_t b[6];
    const int rc = read_back(b, bound.p, sizeof(b), st);
    bound.release();
    if (rc) return rc;
    auto fl = [](uint32_t u) {
        const uint32_t v = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
        float f;
        std::memcpy(&f, &v, 4);
        return f;
    };
    for (int k = 0; k < 3; ++k) { s->bmin[k] = fl(b[k]); s->bmax[k] = fl(b[3 + k]); }
    s->tri_world_stale = true;
    return RRH_OK;
}

int tree_links(rrh_ctx* c, rrh_scene* s, uint32_t top4) {
    const uint32_t n4 = s->nnodes4;
    RRH_TRY(s->parent4.ensure((size_t)n4 * sizeof(uint32_t)));
    k_w4_parent_index<<<grid(n4), kBlock, 0, c->stream>>>((const float4*)s->nodes4.p, n4, (uint32_t*)s->parent4.p);
    RRH_LAUNCH_CHECK();
    s->top4 = std::min(top4, n4);
    return RRH_OK;
}

int refit_bvh_device(rrh_ctx* c, rrh_scene* s) {
    hipStream_t st = c->stream;
    const uint32_t nb = s->nnodes, n4 = s->nnodes4;
    k_tris_refresh<<<grid(s->ntri), kBlock, 0, st>>>((const float*)s->twd.p, s->ntri, (float4*)s->tris.p);
    RRH_LAUNCH_CHECK();
    if (!s->refit_ready) {
        RRH_TRY(s->refit.ensure(((size_t)nb + n4) * 3 * sizeof(uint32_t)));
        uint32_t* r = (uint32_t*)s->refit.p;
        k_bin_parents<<<grid(nb), kBlock, 0, st>>>((const float4*)s->nodes.p, nb, r, r + nb);
        k_w4_parents<<<grid(n4), kBlock, 0, st>>>((const float4*)s->nodes4.p, n4, r + 3 * (size_t)nb, r + 3 * (size_t)nb + n4);
        RRH_LAUNCH_CHECK();
        s->refit_ready = true;
    }
    uint32_t* r = (uint32_t*)s->refit.p;
    uint32_t *bpar = r, *binn = r + nb, *barr = r + 2 * (size_t)nb;
    uint32_t *wpar = r + 3 * (size_t)nb, *winn = wpar + n4, *warr = winn + n4;
    RRH_HIP(hipMemsetAsync(barr, 0, (size_t)nb * 4, st));
    RRH_HIP(hipMemsetAsync(warr, 0, (size_t)n4 * 4, st));
    k_bin_refit<<<grid(nb), kBlock, 0, st>>>((float4*)s->nodes.p, nb, (const float4*)s->tris.p, bpar, binn, barr);
    k_w4_refit<<<grid(n4), kBlock, 0, st>>>((float4*)s->nodes4.p, (uint4*)s->nodesq.p, n4, (const float4*)s->tris.p,
                                            wpar, winn, warr);
    RRH_LAUNCH_CHECK();
    RRH_HIP(hipStreamSynchronize(st));
    return RRH_OK;
}

}  // namespace rrh

// rrh_bvh.cpp -- binned-SAH binary BVH builder (host) and the device node layout.
//
// Replaces the driver-built BLAS/TLAS of the reference (src/rhi/AccelStructure.cpp:
// 98-330, PREFER_FAST_TRACE).  The scene is flattened to world-space triangles
// first, so one BVH covers every instance.
//
// Device layout (64 B per inner node, both child boxes in the parent so one
// node fetch decides both children):
//   float4 a = (L.lo.x, L.hi.x, L.lo.y, L.hi.y)
//   float4 b = (R.lo.x, R.hi.x, R.lo.y, R.hi.y)
//   float4 c = (L.lo.z, L.hi.z, R.lo.z, R.hi.z)
//   int4   d = (L link, R link, 0, 0)   link >= 0: inner node; < 0: leaf ~((first<<3)|(count-1))
// Nodes are emitted depth-first (left child right after its parent).  Triangles
// are stored in leaf order as 3 float4 (v0, v1, v2); v0.w holds the flattened id.
//
// 4-wide layout (128 B per node, one L2 line): the binary tree collapsed by repeatedly
// opening the child of largest surface area (order kept, so the leaves keep their order
// and triangle ranges):
//   float4 [0] lo.x of children 0..3   [1] hi.x   [2] lo.y   [3] hi.y   [4] lo.z   [5] hi.z
//   int4   [6] child links (as above; 0 = empty slot, the root is never a child)   [7] unused
//
// Quantized 4-wide layout (64 B per node, the same tree and links): per axis an origin
// (the node box's lower corner) and a power-of-two step 2^(E-127); each child bound is an
// 8-bit multiple of the step, rounded outward and checked in fp32 exactly as the kernel
// decodes it (origin + q * step, no fused multiply-add), so every decoded child box
// contains the true one and traversal returns the same closest hit:
//   uint [0..2] origin xyz (fp32 bits)   [3] E_x | E_y << 8 | E_z << 16
//   uint [4] lo.x of children 0..3 (one byte each)  [5] hi.x  [6] lo.y  [7] hi.y  [8] lo.z  [9] hi.z
//   int  [10..13] child links   [14..15] unused
#include <algorithm>
#include <cmath>
#include <cstring>
#include <future>
#include <thread>

#include "rrh_internal.h"

namespace rrh {
namespace {

struct Box {
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const Box& b) {
        for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], b.lo[k]); hi[k] = std::max(hi[k], b.hi[k]); }
    }
    void grow_pt(const float* p) {
        for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], p[k]); hi[k] = std::max(hi[k], p[k]); }
    }
    float area() const {
        float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (!(dx >= 0.0f)) return 0.0f;
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

struct TNode {
    Box box;
    int32_t left = -1, right = -1;  // children in the temp array (-1 = leaf)
    uint32_t first = 0, count = 0;
    uint32_t depth = 0;
};

constexpr int kBins = 32;
constexpr uint32_t kMaxLeaf = 8;       // 3-bit count in the link
constexpr uint32_t kMaxDepth = 62;     // device stack holds 64 entries
constexpr uint32_t kMedianDepth = 48;  // below this depth only balanced splits

struct Builder {
    const float* tv;
    uint32_t max_leaf = kMaxLeaf;  // largest leaf (<= kMaxLeaf)
    float trav_cost = 1.0f;        // SAH cost of a traversal step, in triangle tests
    std::vector<Box> tbox;
    std::vector<float> cen;
    std::vector<uint32_t> order;
    std::vector<TNode> nodes;   // filled per subtree then merged
};

// recursive build of order[first, first+count) into `out` (local node list)
void build_rec(Builder& B, uint32_t first, uint32_t count, uint32_t depth, std::vector<TNode>& out,
               uint32_t self) {
    TNode& nd = out[self];
    nd.first = first; nd.count = count; nd.depth = depth;
    Box cb;
    for (uint32_t i = first; i < first + count; ++i) {
        uint32_t t = B.order[i];
        nd.box.grow(B.tbox[t]);
        cb.grow_pt(&B.cen[t * 3]);
    }
    if (count <= 2 && depth < kMaxDepth) {
        if (count <= 2) return;  // leaf
    }
    int best_axis = -1, best_split = -1;
    float best_cost = INFINITY;
    const float leaf_cost = (float)count;
    bool median = depth >= kMedianDepth;
    if (!median) {
        for (int ax = 0; ax < 3; ++ax) {
            float ext = cb.hi[ax] - cb.lo[ax];
            if (!(ext > 0.0f)) continue;
            Box bb[kBins];
            uint32_t bc[kBins] = {0};
            float k = (float)kBins / ext;
            for (uint32_t i = first; i < first + count; ++i) {
                uint32_t t = B.order[i];
                int b = (int)((B.cen[t * 3 + ax] - cb.lo[ax]) * k);
                b = std::min(std::max(b, 0), kBins - 1);
                bc[b]++;
                bb[b].grow(B.tbox[t]);
            }
            float ra[kBins]; uint32_t rc[kBins];
            Box acc; uint32_t n = 0;
            for (int b = kBins - 1; b > 0; --b) {
                acc.grow(bb[b]); n += bc[b];
                ra[b] = acc.area(); rc[b] = n;
            }
            Box lacc; uint32_t ln = 0;
            float inv = 1.0f / std::max(nd.box.area(), 1e-30f);
            for (int b = 0; b < kBins - 1; ++b) {
                lacc.grow(bb[b]); ln += bc[b];
                if (ln == 0 || rc[b + 1] == 0) continue;
                float cost = B.trav_cost + (lacc.area() * (float)ln + ra[b + 1] * (float)rc[b + 1]) * inv;
                if (cost < best_cost) { best_cost = cost; best_axis = ax; best_split = b; }
            }
        }
    }
    uint32_t nl = 0;
    if (!median && best_axis >= 0 && (best_cost < leaf_cost || count > B.max_leaf)) {
        float ext = cb.hi[best_axis] - cb.lo[best_axis];
        float k = (float)kBins / ext;
        auto mid = std::partition(B.order.begin() + first, B.order.begin() + first + count, [&](uint32_t t) {
            int b = (int)((B.cen[t * 3 + best_axis] - cb.lo[best_axis]) * k);
            b = std::min(std::max(b, 0), kBins - 1);
            return b <= best_split;
        });
        nl = (uint32_t)(mid - (B.order.begin() + first));
    } else if (count > B.max_leaf || (median && count > 2)) {
        // object median on the longest centroid axis
        int ax = 0;
        for (int a = 1; a < 3; ++a)
            if (cb.hi[a] - cb.lo[a] > cb.hi[ax] - cb.lo[ax]) ax = a;
        nl = count / 2;
        std::nth_element(B.order.begin() + first, B.order.begin() + first + nl, B.order.begin() + first + count,
                         [&](uint32_t x, uint32_t y) { return B.cen[x * 3 + ax] < B.cen[y * 3 + ax]; });
    } else {
        return;  // leaf
    }
    if (nl == 0 || nl == count) nl = count / 2;
    uint32_t li = (uint32_t)out.size();
    out.emplace_back();
    out.emplace_back();
    out[self].left = (int32_t)li;
    out[self].right = (int32_t)(li + 1);
    // large subtrees in parallel
    if (count > 200000 && depth < 4) {
        std::vector<TNode> rnodes(1);
        auto fut = std::async(std::launch::async, [&]() { build_rec(B, first + nl, count - nl, depth + 1, rnodes, 0); });
        build_rec(B, first, nl, depth + 1, out, li);
        fut.get();
        // splice the right subtree: remap child indices
        uint32_t base = (uint32_t)out.size();
        TNode r0 = rnodes[0];
        for (size_t i = 1; i < rnodes.size(); ++i) {
            TNode t = rnodes[i];
            if (t.left >= 0) { t.left = (int32_t)(base + (uint32_t)t.left - 1); t.right = (int32_t)(base + (uint32_t)t.right - 1); }
            out.push_back(t);
        }
        if (r0.left >= 0) { r0.left = (int32_t)(base + (uint32_t)r0.left - 1); r0.right = (int32_t)(base + (uint32_t)r0.right - 1); }
        out[li + 1] = r0;
    } else {
        build_rec(B, first, nl, depth + 1, out, li);
        build_rec(B, first + nl, count - nl, depth + 1, out, li + 1);
    }
}

int32_t leaf_link(uint32_t first, uint32_t count) { return (int32_t)~((first << 3) | (count - 1u)); }

struct Emitter {
    const Builder& B;
    const std::vector<TNode>& tn;
    BvhOut& out;
    std::vector<int32_t>& leaf_links;  // per temp-tree leaf: its device link
    uint32_t max_depth = 0;
    int32_t child_link(uint32_t c, uint32_t depth) {
        const TNode& t = tn[c];
        if (t.left < 0) {
            uint32_t first = (uint32_t)(out.tris.size() / 3);
            for (uint32_t i = t.first; i < t.first + t.count; ++i) {
                uint32_t id = B.order[i];
                const float* v = B.tv + (size_t)id * 9;
                float4 a, b, c2;
                a.x = v[0]; a.y = v[1]; a.z = v[2]; uint32_t idb = id; std::memcpy(&a.w, &idb, 4);
                b.x = v[3]; b.y = v[4]; b.z = v[5]; b.w = 0.0f;
                c2.x = v[6]; c2.y = v[7]; c2.z = v[8]; c2.w = 0.0f;
                out.tris.push_back(a); out.tris.push_back(b); out.tris.push_back(c2);
            }
            out.leaves++;
            max_depth = std::max(max_depth, depth);
            leaf_links[c] = leaf_link(first, t.count);
            return leaf_links[c];
        }
        return (int32_t)emit(c, depth);
    }
    uint32_t emit(uint32_t n, uint32_t depth) {
        uint32_t idx = (uint32_t)(out.nodes.size() / 4);
        out.nodes.resize(out.nodes.size() + 4);
        const TNode& t = tn[n];
        const Box& L = tn[t.left].box;
        const Box& R = tn[t.right].box;
        float4 a = {L.lo[0], L.hi[0], L.lo[1], L.hi[1]};
        float4 b = {R.lo[0], R.hi[0], R.lo[1], R.hi[1]};
        float4 c = {L.lo[2], L.hi[2], R.lo[2], R.hi[2]};
        int32_t ll = child_link((uint32_t)t.left, depth + 1);
        int32_t rl = child_link((uint32_t)t.right, depth + 1);
        float4 d; std::memcpy(&d.x, &ll, 4); std::memcpy(&d.y, &rl, 4); d.z = 0.0f; d.w = 0.0f;
        out.nodes[idx * 4 + 0] = a; out.nodes[idx * 4 + 1] = b;
        out.nodes[idx * 4 + 2] = c; out.nodes[idx * 4 + 3] = d;
        return idx;
    }
};

// one axis of a quantized node: biased exponent E and the children's 8-bit bounds
void quantize_axis(float origin, float top, const float* lo, const float* hi, uint32_t nch, uint32_t& E,
                   uint8_t* qlo, uint8_t* qhi) {
    const double ext = (double)top - (double)origin;
    int e = -126;  // smallest normal step
    while (e < 127 && std::ldexp(255.0, e) < ext) ++e;
    for (;; ++e) {
        const float step = std::ldexp(1.0f, e);
        bool ok = true;
        for (uint32_t i = 0; i < nch && ok; ++i) {
            double fl = std::floor(((double)lo[i] - (double)origin) / (double)step);
            double fh = std::ceil(((double)hi[i] - (double)origin) / (double)step);
            int a = (int)std::max(0.0, std::min(255.0, fl)), b = (int)std::max(0.0, std::min(255.0, fh));
            // the kernel's decode, in fp32: move outward until the decoded box contains the child's
            volatile float dl = origin + (float)a * step, dh = origin + (float)b * step;
            while (a > 0 && dl > lo[i]) { --a; dl = origin + (float)a * step; }
            while (b < 255 && dh < hi[i]) { ++b; dh = origin + (float)b * step; }
            if (dl > lo[i] || dh < hi[i]) ok = false;
            qlo[i] = (uint8_t)a;
            qhi[i] = (uint8_t)b;
        }
        if (ok || e >= 127) break;
    }
    E = (uint32_t)(e + 127);
    for (uint32_t i = nch; i < 4; ++i) { qlo[i] = 0; qhi[i] = 0; }
}

// collapse of the temp tree into 4-wide nodes (after the binary emission has assigned
// every leaf its triangle range)
struct Emitter4 {
    const std::vector<TNode>& tn;
    const std::vector<int32_t>& leaf_links;
    BvhOut& out;
    uint32_t max_depth = 0;
    // returns the node index; `stack` receives the stack entries the subtree can need
    uint32_t emit(uint32_t n, uint32_t depth, uint32_t& stack) {
        uint32_t ch[4] = {(uint32_t)tn[n].left, (uint32_t)tn[n].right, 0, 0};
        uint32_t nch = 2;
        while (nch < 4) {  // open the inner child of largest surface area, in place
            int best = -1;
            float best_a = -1.0f;
            for (uint32_t i = 0; i < nch; ++i)
                if (tn[ch[i]].left >= 0 && tn[ch[i]].box.area() > best_a) { best = (int)i; best_a = tn[ch[i]].box.area(); }
            if (best < 0) break;
            const uint32_t c = ch[best];
            for (int j = (int)nch; j > best + 1; --j) ch[j] = ch[j - 1];
            ch[best] = (uint32_t)tn[c].left;
            ch[best + 1] = (uint32_t)tn[c].right;
            ++nch;
        }
        const uint32_t idx = (uint32_t)(out.nodes4.size() / 8);
        out.nodes4.resize(out.nodes4.size() + 8);
        int32_t links[4] = {0, 0, 0, 0};
        float lo[3][4] = {}, hi[3][4] = {};
        uint32_t sub = 0;
        for (uint32_t i = 0; i < nch; ++i) {
            const TNode& t = tn[ch[i]];
            for (int k = 0; k < 3; ++k) { lo[k][i] = t.box.lo[k]; hi[k][i] = t.box.hi[k]; }
            if (t.left < 0) {
                links[i] = leaf_links[ch[i]];
                max_depth = std::max(max_depth, depth + 1);
            } else {
                uint32_t st = 0;
                links[i] = (int32_t)emit(ch[i], depth + 1, st);
                sub = std::max(sub, st);
            }
        }
        stack = (nch - 1) + sub;  // a visit pushes at most nch - 1 siblings, then descends
        float4* o = out.nodes4.data() + (size_t)idx * 8;
        for (int k = 0; k < 3; ++k) {
            o[2 * k] = make_float4(lo[k][0], lo[k][1], lo[k][2], lo[k][3]);
            o[2 * k + 1] = make_float4(hi[k][0], hi[k][1], hi[k][2], hi[k][3]);
        }
        std::memcpy(&o[6], links, 16);
        o[7] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        // quantized copy (node index idx, 16 words)
        out.nodesq.resize(out.nodes4.size() / 2);
        uint32_t w[16] = {};
        uint32_t E[3];
        uint8_t ql[3][4], qh[3][4];
        for (int k = 0; k < 3; ++k) {
            float org = lo[k][0], top = hi[k][0];
            for (uint32_t i = 1; i < nch; ++i) { org = std::min(org, lo[k][i]); top = std::max(top, hi[k][i]); }
            quantize_axis(org, top, lo[k], hi[k], nch, E[k], ql[k], qh[k]);
            std::memcpy(&w[k], &org, 4);
            w[4 + 2 * k] = ql[k][0] | (ql[k][1] << 8) | (ql[k][2] << 16) | ((uint32_t)ql[k][3] << 24);
            w[5 + 2 * k] = qh[k][0] | (qh[k][1] << 8) | (qh[k][2] << 16) | ((uint32_t)qh[k][3] << 24);
        }
        w[3] = E[0] | (E[1] << 8) | (E[2] << 16);
        std::memcpy(&w[10], links, 16);
        std::memcpy(out.nodesq.data() + (size_t)idx * 4, w, 64);
        return idx;
    }
};

// Renumbers the 4-wide nodes so that the first `top` nodes of a breadth-first walk (children
// in slot order) come first, and the rest keep their depth-first emission order.  Every
// prefix of the result then holds the ancestors of its nodes (the top of the tree, which the
// traversal kernels can stage in LDS), and links still point forward.  Closest hits do not
// depend on node order.
void reorder_top_bfs(BvhOut& out, uint32_t top) {
    const uint32_t n4 = (uint32_t)(out.nodes4.size() / 8);
    auto links = [&](uint32_t n, int32_t l[4]) { std::memcpy(l, &out.nodes4[(size_t)n * 8 + 6], 16); };
    std::vector<uint32_t> order{0};
    for (size_t i = 0; i < order.size() && order.size() < top; ++i) {
        int32_t l[4];
        links(order[i], l);
        for (int s = 0; s < 4 && order.size() < top; ++s)
            if (l[s] > 0) order.push_back((uint32_t)l[s]);
    }
    std::vector<uint32_t> id(n4, ~0u);
    for (uint32_t i = 0; i < (uint32_t)order.size(); ++i) id[order[i]] = i;
    uint32_t next = (uint32_t)order.size();
    for (uint32_t n = 0; n < n4; ++n)
        if (id[n] == ~0u) id[n] = next++;
    std::vector<float4> n4o(out.nodes4.size()), nqo(out.nodesq.size());
    for (uint32_t n = 0; n < n4; ++n) {
        int32_t l[4];
        links(n, l);
        for (int s = 0; s < 4; ++s)
            if (l[s] > 0) l[s] = (int32_t)id[l[s]];
        std::memcpy(&n4o[(size_t)id[n] * 8], &out.nodes4[(size_t)n * 8], 128);
        std::memcpy(&n4o[(size_t)id[n] * 8 + 6], l, 16);
        std::memcpy(&nqo[(size_t)id[n] * 4], &out.nodesq[(size_t)n * 4], 64);
        std::memcpy(reinterpret_cast<uint32_t*>(&nqo[(size_t)id[n] * 4]) + 10, l, 16);
    }
    out.nodes4.swap(n4o);
    out.nodesq.swap(nqo);
}

}  // namespace

bool build_bvh(const float* tri_v, uint32_t ntri, BvhOut& out, std::string& err, uint32_t max_leaf,
               float trav_cost) {
    out = BvhOut();
    if (max_leaf < 1 || max_leaf > kMaxLeaf || !(trav_cost > 0.0f)) { err = "bad BVH build parameters"; return false; }
    if (ntri == 0) { err = "scene has no triangles"; return false; }
    if (ntri >= (1u << 28)) { err = "too many triangles for the 28-bit leaf index"; return false; }
    Builder B;
    B.tv = tri_v;
    B.max_leaf = max_leaf;
    B.trav_cost = trav_cost;
    B.tbox.resize(ntri);
    B.cen.resize((size_t)ntri * 3);
    B.order.resize(ntri);
    for (uint32_t i = 0; i < ntri; ++i) {
        Box b;
        for (int c = 0; c < 3; ++c) b.grow_pt(tri_v + (size_t)i * 9 + c * 3);
        B.tbox[i] = b;
        for (int k = 0; k < 3; ++k) B.cen[(size_t)i * 3 + k] = 0.5f * (b.lo[k] + b.hi[k]);
        B.order[i] = i;
    }
    std::vector<TNode> tn(1);
    tn.reserve((size_t)ntri * 2 / 3 + 16);
    build_rec(B, 0, ntri, 0, tn, 0);
    std::vector<int32_t> leaf_links(tn.size() + 1, 0);
    Emitter E{B, tn, out, leaf_links};
    if (tn[0].left < 0) {
        // single leaf: a root whose two children are that same leaf
        uint32_t n = tn[0].count;
        if (n > kMaxLeaf) { err = "internal: oversized root leaf"; return false; }
        TNode leaf = tn[0];
        tn.push_back(leaf);
        tn[0].left = 1; tn[0].right = 1;
        out.nodes.resize(4);
        const Box& L = tn[1].box;
        int32_t link = E.child_link(1, 1);
        float4 a = {L.lo[0], L.hi[0], L.lo[1], L.hi[1]};
        float4 c = {L.lo[2], L.hi[2], L.lo[2], L.hi[2]};
        float4 d; std::memcpy(&d.x, &link, 4); std::memcpy(&d.y, &link, 4); d.z = 0.0f; d.w = 0.0f;
        out.nodes[0] = a; out.nodes[1] = a; out.nodes[2] = c; out.nodes[3] = d;
    } else {
        E.emit(0, 0);
    }
    out.depth = E.max_depth;
    if (out.depth > kMaxDepth + 1) { err = "BVH deeper than the traversal stack"; return false; }
    // 4-wide collapse (the single-leaf root: one leaf child, three empty slots)
    Emitter4 E4{tn, leaf_links, out};
    if (tn[0].left < 0 || (tn[0].left == 1 && tn[0].right == 1 && tn.size() == 2)) {
        out.nodes4.assign(8, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        const Box& L = tn.back().box;
        out.nodes4[0] = make_float4(L.lo[0], 0, 0, 0); out.nodes4[1] = make_float4(L.hi[0], 0, 0, 0);
        out.nodes4[2] = make_float4(L.lo[1], 0, 0, 0); out.nodes4[3] = make_float4(L.hi[1], 0, 0, 0);
        out.nodes4[4] = make_float4(L.lo[2], 0, 0, 0); out.nodes4[5] = make_float4(L.hi[2], 0, 0, 0);
        int32_t links[4] = {leaf_links[tn.size() - 1], 0, 0, 0};
        std::memcpy(&out.nodes4[6], links, 16);
        out.nodesq.assign(4, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        uint32_t w[16] = {};
        for (int k = 0; k < 3; ++k) {
            uint32_t E;
            uint8_t ql[4], qh[4];
            quantize_axis(L.lo[k], L.hi[k], &L.lo[k], &L.hi[k], 1, E, ql, qh);
            std::memcpy(&w[k], &L.lo[k], 4);
            w[3] |= E << (8 * k);
            w[4 + 2 * k] = ql[0];
            w[5 + 2 * k] = qh[0];
        }
        std::memcpy(&w[10], links, 16);
        std::memcpy(out.nodesq.data(), w, 64);
        out.depth4 = 1;
        out.stack4 = 0;
    } else {
        uint32_t st = 0;
        E4.emit(0, 0, st);
        out.depth4 = E4.max_depth;
        out.stack4 = st;
        reorder_top_bfs(out, kTopBfsNodes);
    }
    if (out.stack4 > 64) { err = "4-wide BVH needs more than 64 traversal stack entries"; return false; }
    // structural check before anything reaches the device: links in range, every
    // triangle referenced exactly once (the single-leaf root references its leaf twice)
    const uint32_t nn = (uint32_t)(out.nodes.size() / 4), ntris = (uint32_t)(out.tris.size() / 3);
    if (ntris != ntri) { err = "internal: triangle count changed during the build"; return false; }
    std::vector<uint8_t> seen(ntri, 0);
    for (uint32_t n = 0; n < nn; ++n) {
        int32_t links[2];
        std::memcpy(&links[0], &out.nodes[n * 4 + 3].x, 4);
        std::memcpy(&links[1], &out.nodes[n * 4 + 3].y, 4);
        for (int k = 0; k < 2; ++k) {
            int32_t l = links[k];
            if (l >= 0) {
                if ((uint32_t)l >= nn || (uint32_t)l <= n) { err = "internal: bad inner link"; return false; }
            } else {
                uint32_t first = ((uint32_t)~l) >> 3, cnt = (((uint32_t)~l) & 7u) + 1u;
                if ((uint64_t)first + cnt > ntris) { err = "internal: leaf out of range"; return false; }
                if (nn == 1 && k == 1 && links[0] == links[1]) continue;
                for (uint32_t t = first; t < first + cnt; ++t) {
                    if (seen[t]) { err = "internal: triangle referenced twice"; return false; }
                    seen[t] = 1;
                }
            }
        }
    }
    for (uint32_t t = 0; t < ntri; ++t)
        if (!seen[t]) { err = "internal: triangle not referenced"; return false; }
    // the same checks for the 4-wide nodes
    const uint32_t n4 = (uint32_t)(out.nodes4.size() / 8);
    std::fill(seen.begin(), seen.end(), 0);
    for (uint32_t n = 0; n < n4; ++n) {
        int32_t links[4];
        std::memcpy(links, &out.nodes4[(size_t)n * 8 + 6], 16);
        for (int k = 0; k < 4; ++k) {
            const int32_t l = links[k];
            if (l == 0) continue;
            if (l > 0) {
                if ((uint32_t)l >= n4 || (uint32_t)l <= n) { err = "internal: bad 4-wide inner link"; return false; }
            } else {
                uint32_t first = ((uint32_t)~l) >> 3, cnt = (((uint32_t)~l) & 7u) + 1u;
                if ((uint64_t)first + cnt > ntris) { err = "internal: 4-wide leaf out of range"; return false; }
                for (uint32_t t = first; t < first + cnt; ++t) {
                    if (seen[t]) { err = "internal: triangle referenced twice (4-wide)"; return false; }
                    seen[t] = 1;
                }
            }
        }
    }
    for (uint32_t t = 0; t < ntri; ++t)
        if (!seen[t]) { err = "internal: triangle not referenced (4-wide)"; return false; }
    // quantized nodes: same links, and every decoded child box contains the fp32 one
    if (out.nodesq.size() != (size_t)n4 * 4) { err = "internal: quantized node count"; return false; }
    for (uint32_t n = 0; n < n4; ++n) {
        uint32_t w[16];
        std::memcpy(w, out.nodesq.data() + (size_t)n * 4, 64);
        if (std::memcmp(&w[10], &out.nodes4[(size_t)n * 8 + 6], 16) != 0) { err = "internal: quantized links"; return false; }
        const float* f = reinterpret_cast<const float*>(out.nodes4.data() + (size_t)n * 8);
        for (int k = 0; k < 3; ++k) {
            float org, step;
            std::memcpy(&org, &w[k], 4);
            const uint32_t sb = ((w[3] >> (8 * k)) & 255u) << 23;
            std::memcpy(&step, &sb, 4);
            for (int i = 0; i < 4; ++i) {
                int32_t l;
                std::memcpy(&l, &w[10 + i], 4);
                if (l == 0) continue;
                const float dl = org + (float)((w[4 + 2 * k] >> (8 * i)) & 255u) * step;
                const float dh = org + (float)((w[5 + 2 * k] >> (8 * i)) & 255u) * step;
                if (!(dl <= f[(2 * k) * 4 + i] && dh >= f[(2 * k + 1) * 4 + i])) {
                    err = "internal: quantized box does not contain the child";
                    return false;
                }
            }
        }
    }
    return true;
}

}  // namespace rrh

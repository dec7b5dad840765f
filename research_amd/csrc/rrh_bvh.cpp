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
                float cost = 1.0f + (lacc.area() * (float)ln + ra[b + 1] * (float)rc[b + 1]) * inv;
                if (cost < best_cost) { best_cost = cost; best_axis = ax; best_split = b; }
            }
        }
    }
    uint32_t nl = 0;
    if (!median && best_axis >= 0 && (best_cost < leaf_cost || count > kMaxLeaf)) {
        float ext = cb.hi[best_axis] - cb.lo[best_axis];
        float k = (float)kBins / ext;
        auto mid = std::partition(B.order.begin() + first, B.order.begin() + first + count, [&](uint32_t t) {
            int b = (int)((B.cen[t * 3 + best_axis] - cb.lo[best_axis]) * k);
            b = std::min(std::max(b, 0), kBins - 1);
            return b <= best_split;
        });
        nl = (uint32_t)(mid - (B.order.begin() + first));
    } else if (count > kMaxLeaf || (median && count > 2)) {
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
            return leaf_link(first, t.count);
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

}  // namespace

bool build_bvh(const float* tri_v, uint32_t ntri, BvhOut& out, std::string& err) {
    out = BvhOut();
    if (ntri == 0) { err = "scene has no triangles"; return false; }
    if (ntri >= (1u << 28)) { err = "too many triangles for the 28-bit leaf index"; return false; }
    Builder B;
    B.tv = tri_v;
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
    Emitter E{B, tn, out};
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
    return true;
}

}  // namespace rrh

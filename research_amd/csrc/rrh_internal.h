// rrh_internal.h -- host-side objects behind the opaque C handles.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/rrh.h"
#include "rrh_device.h"

namespace rrh {

void set_error(const std::string& msg);

#define RRH_HIP(call)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            ::rrh::set_error(std::string(#call) + ": " + hipGetErrorString(e_));            \
            return RRH_E_HIP;                                                               \
        }                                                                                   \
    } while (0)
#define RRH_ARG(cond, msg)                                                                  \
    do {                                                                                    \
        if (!(cond)) { ::rrh::set_error(msg); return RRH_E_ARG; }                           \
    } while (0)
#define RRH_LAUNCH_CHECK()                                                                  \
    do {                                                                                    \
        hipError_t e_ = hipGetLastError();                                                  \
        if (e_ != hipSuccess) {                                                             \
            ::rrh::set_error(std::string("kernel launch: ") + hipGetErrorString(e_));        \
            return RRH_E_HIP;                                                               \
        }                                                                                   \
    } while (0)

// device buffer that grows on demand
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    int ensure(size_t n);
    void release();
};

// BVH produced by the builder (rrh_bvh.cpp): the binary SAH tree and its 4-wide collapse,
// which visit the leaves in the same order and so share one triangle array
struct BvhOut {
    std::vector<float4> nodes;   // 4 per binary inner node
    std::vector<float4> nodes4;  // 8 per 4-wide node
    std::vector<float4> nodesq;  // 4 per quantized 4-wide node (16 words, see rrh_bvh.cpp)
    std::vector<float4> tris;    // 3 per triangle, leaf order
    uint32_t depth = 0, leaves = 0;
    uint32_t depth4 = 0, stack4 = 0;  // 4-wide: depth and the traversal stack entries it can need
};
// the host builder numbers the first kTopBfsNodes 4-wide nodes breadth-first (the device
// builder numbers all of them so): any prefix of that many nodes is the top of the tree
constexpr uint32_t kTopBfsNodes = 1024;
// tri_v: ntri x 9 floats (world space).  Returns false with a message on failure.
// max_leaf: largest leaf (1..8); trav_cost: SAH cost of a traversal step relative to a triangle test
bool build_bvh(const float* tri_v, uint32_t ntri, BvhOut& out, std::string& err, uint32_t max_leaf = 8,
               float trav_cost = 1.0f);

// BVH build on the device (rrh_bvh_gpu.hip) from the world triangles tw (device, ntri x 9
// floats): fills s->nodes, nodes4, nodesq, tris and the tree statistics; RRH_OK or an error
// code with the message set
int build_bvh_device(rrh_ctx* c, rrh_scene* s, const float* tw, uint32_t max_leaf, float trav_cost, uint32_t radius);
// world triangles of the scene's current instances into s->twd, triangle-light records and
// bounds updated (rrh_bvh_gpu.hip)
int flatten_device(rrh_ctx* c, rrh_scene* s);
// every box of the three node formats recomputed bottom-up from s->twd, topology kept
int refit_bvh_device(rrh_ctx* c, rrh_scene* s);
// after a build: the parent of every 4-wide node (s->parent4) and the breadth-first prefix
// length s->top4 = min(top4, node count)
int tree_links(rrh_ctx* c, rrh_scene* s, uint32_t top4);
// this device's address of rrh_pt_render_noalloc's scratch array (device must be current)
int pt_scratch_init(int device, void** out);

// procedural scene generators (rrh_scenegen.cpp)
struct SceneData {
    std::vector<rrh_vertex> verts;
    std::vector<uint32_t> idx;
    std::vector<rrh_mesh> meshes;
    std::vector<rrh_material> mats;
    std::vector<rrh_instance> insts;
    std::vector<rrh_light> lights;
    uint32_t flags = 0;
    float eye[3] = {0, 0, 3}, center[3] = {0, 0, 0}, fov = 45.0f;
};
bool generate_scene(int kind, uint32_t seed, float scale, SceneData& out, std::string& err);

}  // namespace rrh

struct rrh_ctx {
    int device = 0;
    int cu_count = 256;            // compute units of the device (persistent grids)
    bool persistent = false;       // RRH_OPT_PERSISTENT: ray queues traced by persistent waves
    uint32_t persistent_waves = 1; // RRH_OPT_PERSISTENT's value: 1 = as many waves as the CUs hold, N > 1 caps them
    uint32_t refill = 16;          // RRH_OPT_REFILL: idle lanes at which a persistent wave takes new rays
    uint32_t suspend_visits = 0;   // RRH_OPT_SUSPEND_VISITS: node visits after which a ray is resumed later (0: off)
    uint32_t rpt_step_visits = 0;  // RRH_OPT_RPT_STEP_VISITS: ReSTIR PT initial stage, node visits per traversal pass
    rrh::DevBuf susp;              // continuation records of suspended rays
    bool speculative = true;       // RRH_OPT_SPECULATIVE: traversal parks leaves (default: measured faster)
    uint32_t bvh_max_leaf = 8;     // RRH_OPT_BVH_MAX_LEAF: applies to later scene uploads
    uint32_t bvh_trav_cost = 100;  // RRH_OPT_BVH_TRAV_COST: SAH traversal cost x 100 (later uploads)
    int bvh_builder = RRH_BVH_BUILD_HOST;  // RRH_OPT_BVH_BUILDER (later uploads)
    uint32_t ploc_radius = 16;     // RRH_OPT_PLOC_RADIUS: device builder's neighbour window
    uint32_t ray_sort_bits = 0;    // RRH_OPT_RAY_SORT: 0 = extension rays traced in queue order
    rrh::DevBuf sort_keys, sort_perm, sort_tmp;  // ray sort: keys / slots (x2 each), rocPRIM scratch
    hipStream_t stream = nullptr;
    hipEvent_t markers[16] = {};
    hipEvent_t join_event = nullptr;  // rrh_ctx_join: recorded on this context's stream
    // path-tracer wavefront state (sized for the largest frame seen)
    rrh::DevBuf q_o[2], q_d[2], q_thp[2];
    rrh::DevBuf hits, sh_o, sh_d, sh_c;
    rrh::DevBuf counters;         // per-bounce queue counts (see rrh_pt.hip)
    rrh::DevBuf vtx;              // recorded path vertices (guiding)
    rrh::DevBuf train;            // finalised training samples (4 float4 each)
    rrh::DevBuf train_cnt;        // device counter of training samples
    rrh::DevBuf fin_chunks, fin_tmp;  // record-order compaction: per-chunk counts + offsets, scan scratch
    uint32_t train_capacity = 0;
    uint32_t stamp = 0;           // vertex-record generation
    uint32_t last_rec_bounces = 0;
    uint32_t last_max_bounces = 0;
    const uint32_t* last_counters = nullptr;  // per-bounce counters of the last path-traced frame
    void* noalloc_scratch = nullptr;         // this device's rrh_pt_render_noalloc scratch (hipGetSymbolAddress)
    uint32_t last_record_stride = 0;
    uint32_t last_npix = 0;
    rrh::DevBuf di_counter;
    rrh::DevBuf di_pack;          // ReSTIR DI spatial pass: one 16-byte neighbour record per pixel
    rrh::DevBuf rpt_counter;      // ReSTIR PT ray counters (see rrh_restir_pt.hip)
    rrh::DevBuf rgi_counter;      // ReSTIR GI ray counters (see rrh_restir_gi.hip)
    rrh::DevBuf totals;           // uint64 x 3: primary, extension, shadow (device-side sums)
    rrh::DevBuf fetch;            // uint64 x 3: node fetches, tri fetches, rays (instrumented traversal)
    bool profiling = false;
    int stack_kind = RRH_STACK_LDS32;
    uint32_t top_nodes = 0;        // RRH_OPT_TRAV_TOP_NODES: top-of-tree nodes staged in LDS
    bool trace_pairs = true;       // RRH_OPT_TRACE_PAIRS: extension rays of bounce b with shadow rays of b - 1
    bool xcd_remap = false;
    bool restir_cull = false;      // RRH_OPT_RESTIR_CULL: ReSTIR PT / GI rays cull back faces
    int bvh_nodes = RRH_NODES_WIDE4Q;  // RRH_OPT_BVH_NODES (quantized 4-wide: fastest measured)
    std::vector<hipEvent_t> prof_events;   // pairs around each traversal launch
    uint32_t prof_used = 0;
    hipEvent_t prof_mark = nullptr;       // rrh_profile_mark: time origin of rrh_profile_spans
    rrh::DevBuf scratch;          // small general-purpose scratch
    rrh::DevBuf mlp_prep;         // learned guide: per queue slot (pos, hit flag) (normal) (wo)
};

struct rrh_scene {
    rrh_ctx* ctx = nullptr;
    rrh::DevBuf nodes, nodes4, nodesq, tris, tri_ref, tri_light, tl, verts, idx, meshes, mats, insts, lights, cam;
    rrh::DevBuf tri_vtx;          // per flattened id: global indices of its three vertices, instance
    uint32_t ntri = 0, nnodes = 0, nl = 0, ntl = 0, flags = 0, depth = 0, leaves = 0;
    uint32_t nnodes4 = 0, depth4 = 0, stack4 = 0;
    rrh::DevBuf parent4;          // parent of each 4-wide node (the stackless traversal climbs them)
    uint32_t top4 = 0;             // leading 4-wide nodes in breadth-first order (see kTopBfsNodes)
    int node_format() const;       // the format the context traverses (RRH_NODES_*)
    uint32_t stack_need() const;   // traversal stack entries the traversed tree can need
    float build_ms = 0.0f;
    uint32_t build_passes = 0;     // device builder: clustering passes
    uint32_t ninst = 0;
    std::vector<uint32_t> inst_mesh;  // mesh index per instance (an update must keep them)
    rrh::DevBuf twd;               // world triangles on the device (device builds, instance updates)
    rrh::DevBuf refit;             // refit: parent links, inner-child counts, arrival counters
    bool refit_ready = false;      // refit links describe the current tree
    mutable bool tri_world_stale = false;  // host copy behind the device one (after an update)
    float bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};
    mutable std::vector<float> tri_world;  // flattened world triangles (host copy, 9 floats each)
    rrh_camera cam_host{};
    bool cam_set = false;
    rrh::DevScene dev() const;
};

// guiding grid (rrh_guide.hip): the reference's hash grid or the 8-lobe EM grid
struct rrh_grid {
    rrh_ctx* ctx = nullptr;
    rrh_grid_params p{};
    uint32_t ncell = 0;
    rrh::DevBuf accum, fitted;              // HASH
    rrh::DevBuf lobes, stats, fstats, valid; // VMM
    rrh::DevBuf em_keys, em_vals, em_tmp, em_start;  // VMM E-step: sorted (cell, sample) pairs, cell ranges
};

struct rrh_image {
    rrh_ctx* ctx = nullptr;
    uint32_t w = 0, h = 0;
    int fmt = 0;
    size_t bytes = 0;
    void* p = nullptr;
};

namespace rrh {
size_t format_bytes(int fmt);
// stack kind a launch uses for a scene of the given BVH depth
// (need = entries the traversal can push: the binary depth, or the 4-wide bound); a tree
// that may outgrow 32 LDS entries gets 16 in LDS plus a private overflow (measured faster
// than 32 + overflow: the smaller LDS footprint keeps more waves resident)
// (RRH_STACK_NONE: only the path tracer's queue kernels run stackless; the rest get the
// LDS stack that kind would give)
inline int stack_for(const rrh_ctx* c, uint32_t need) {
    if ((c->stack_kind == RRH_STACK_LDS32 || c->stack_kind == RRH_STACK_NONE) && need > 32) return RRH_STACK_HYBRID16;
    if (c->stack_kind == RRH_STACK_NONE) return RRH_STACK_LDS32;
    return c->stack_kind;
}
}

/*
 * rrh.h -- C ABI of librrh.so, the MI355X (gfx950) implementation of the
 * reference renderer's hot path: reference path tracer + accumulation,
 * ReSTIR DI, edge-avoiding A-trous, vMF path guiding (reference single-lobe
 * hash grid and an 8-lobe EM-fitted dense grid) and a learned MLP guide.
 *
 * Conventions
 *  - Every entry point returns int: RRH_OK (0) or a negative RRH_E_* code.
 *    rrh_last_error() returns a thread-local description of the last failure.
 *    No C++ exception crosses this boundary.
 *  - Host pointers are owned by the caller; device memory is owned by the
 *    library and released by the matching rrh_*_destroy.  Each context owns
 *    one HIP stream; all work of a context is ordered on it.  One thread per
 *    context.
 *  - Scene records mirror the reference's GpuScene.h byte for byte.
 *
 * Each declaration cites the reference interface it replaces
 * (paths relative to the reference repository root).
 */
#ifndef RRH_H
#define RRH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RRH_OK 0
#define RRH_E_ARG (-1)       /* invalid argument (null pointer, size mismatch, bad enum) */
#define RRH_E_HIP (-2)       /* a HIP runtime call failed */
#define RRH_E_STATE (-3)     /* object not in a usable state (e.g. scene with no triangles) */
#define RRH_E_NOMEM (-4)     /* device or host allocation failed */
#define RRH_E_RANGE (-5)     /* size above a documented limit */

/* ---- scene records: src/scene/GpuScene.h:17-114 ------------------------- */
typedef struct { float pos[3]; float normal[3]; float uv[2]; float tangent[4]; } rrh_vertex;      /* 48 B, GpuVertex */
typedef struct { uint32_t vertex_base, vertex_count, index_byte_offset, index_count,
                 material_index, _pad[3]; } rrh_mesh;                                              /* 32 B, GpuMesh */
typedef struct { float base_color[4]; float metallic, roughness; float emissive[3]; float ior;
                 uint32_t albedo_tex_idx, normal_tex_idx, metallic_roughness_tex_idx,
                          emissive_tex_idx, sampler_idx;
                 uint32_t flags; /* GpuMaterial._pad[0]: RRH_MAT_* */
                 uint32_t _pad1, _pad2; } rrh_material;                                            /* 72 B, GpuMaterial */
typedef struct { float transform[16]; /* row-major */ uint32_t mesh_index; uint32_t _pad[3]; } rrh_instance; /* 80 B */
typedef struct { float position[3]; uint32_t type; /* 0 point, 1 directional */
                 float direction[3]; float radius; float emission[3]; float intensity; } rrh_light; /* 48 B */
typedef struct { float view[16], inv_view[16], inv_proj[16], prev_view_proj[16];
                 float position[3]; float near_plane; float direction[3]; float far_plane;
                 uint32_t frame_index, screen_width, screen_height, _pad; } rrh_camera;            /* 304 B, GpuCameraData */

/* compile-time layout checks: sizes and field offsets of src/scene/GpuScene.h:17-98, so a
 * caller built against this header fails to compile if a record drifts from the reference */
#if defined(__cplusplus)
#define RRH_STATIC_ASSERT(c, m) static_assert(c, m)
#elif defined(__STDC_VERSION__) && __STDC_VERSION__ >= 201112L
#define RRH_STATIC_ASSERT(c, m) _Static_assert(c, m)
#else
#define RRH_STATIC_ASSERT(c, m)
#endif
RRH_STATIC_ASSERT(sizeof(rrh_vertex) == 48, "rrh_vertex must be 48 B (GpuVertex)");
RRH_STATIC_ASSERT(offsetof(rrh_vertex, normal) == 12 && offsetof(rrh_vertex, uv) == 24 &&
                  offsetof(rrh_vertex, tangent) == 32, "rrh_vertex field offsets (GpuVertex)");
RRH_STATIC_ASSERT(sizeof(rrh_mesh) == 32, "rrh_mesh must be 32 B (GpuMesh)");
RRH_STATIC_ASSERT(offsetof(rrh_mesh, index_byte_offset) == 8 && offsetof(rrh_mesh, material_index) == 16,
                  "rrh_mesh field offsets (GpuMesh)");
RRH_STATIC_ASSERT(sizeof(rrh_material) == 72, "rrh_material must be 72 B (GpuMaterial)");
RRH_STATIC_ASSERT(offsetof(rrh_material, metallic) == 16 && offsetof(rrh_material, emissive) == 24 &&
                  offsetof(rrh_material, ior) == 36 && offsetof(rrh_material, albedo_tex_idx) == 40 &&
                  offsetof(rrh_material, sampler_idx) == 56 && offsetof(rrh_material, flags) == 60,
                  "rrh_material field offsets (GpuMaterial)");
RRH_STATIC_ASSERT(sizeof(rrh_instance) == 80, "rrh_instance must be 80 B (GpuInstance)");
RRH_STATIC_ASSERT(offsetof(rrh_instance, mesh_index) == 64, "rrh_instance field offsets (GpuInstance)");
RRH_STATIC_ASSERT(sizeof(rrh_light) == 48, "rrh_light must be 48 B (GpuLight)");
RRH_STATIC_ASSERT(offsetof(rrh_light, type) == 12 && offsetof(rrh_light, direction) == 16 &&
                  offsetof(rrh_light, radius) == 28 && offsetof(rrh_light, emission) == 32 &&
                  offsetof(rrh_light, intensity) == 44, "rrh_light field offsets (GpuLight)");
RRH_STATIC_ASSERT(sizeof(rrh_camera) == 304, "rrh_camera must be 304 B (GpuCameraData)");
RRH_STATIC_ASSERT(offsetof(rrh_camera, inv_view) == 64 && offsetof(rrh_camera, inv_proj) == 128 &&
                  offsetof(rrh_camera, prev_view_proj) == 192 && offsetof(rrh_camera, position) == 256 &&
                  offsetof(rrh_camera, near_plane) == 268 && offsetof(rrh_camera, direction) == 272 &&
                  offsetof(rrh_camera, far_plane) == 284 && offsetof(rrh_camera, frame_index) == 288 &&
                  offsetof(rrh_camera, screen_width) == 292 && offsetof(rrh_camera, screen_height) == 296,
                  "rrh_camera field offsets (GpuCameraData)");

#define RRH_MAT_PURE_LAMBERT 1u   /* f = albedo/pi, cosine sampling (no GGX lobe) */
#define RRH_SCENE_TRI_LIGHTS 1u   /* emissive triangles become NEE / ReSTIR lights, MIS on BSDF hits */

/* ---- opaque handles -------------------------------------------------------- */
typedef struct rrh_ctx rrh_ctx;
typedef struct rrh_scene rrh_scene;
typedef struct rrh_image rrh_image;
typedef struct rrh_grid rrh_grid;     /* guiding grid (reference hash grid or EM grid) */
typedef struct rrh_mlp rrh_mlp;       /* learned guide */

/* ---- context: replaces rr::rhi::Device + the frame's command recorder
 *      (src/rhi/Device.cpp:300-369 one_time_submit, :270-298 submit_frame) -- */
int rrh_ctx_create(int device, rrh_ctx** out);
int rrh_ctx_destroy(rrh_ctx* ctx);
int rrh_ctx_synchronize(rrh_ctx* ctx);
/* no reference counterpart (the reference records every pass into one queue): work enqueued on
 * `waiter` after this call starts only after everything enqueued on `signaler` before it has
 * finished (an event recorded on the signaler's stream, waited for on the waiter's; the host
 * does not block).  Both contexts on one device. */
int rrh_ctx_join(rrh_ctx* waiter, rrh_ctx* signaler);
const char* rrh_last_error(void);
/* device properties used for roofline arithmetic: CU count, clock, name */
int rrh_device_info(rrh_ctx* ctx, int* cu_count, int* clock_khz, char* name, size_t name_len);
/* elapsed ms between two markers recorded on the context stream */
int rrh_marker_record(rrh_ctx* ctx, int slot);             /* slot in [0, 16) */
int rrh_marker_elapsed(rrh_ctx* ctx, int a, int b, float* ms);

/* ---- scene: replaces Scene::upload (src/scene/Scene.cpp:255-515) and the
 *      BLAS/TLAS build (src/rhi/AccelStructure.cpp:98-330) with a flattened
 *      world-space triangle list and a binned-SAH binary BVH ------------- */
typedef struct {
    uint32_t tri_count, node_count, bvh_depth, leaf_count;
    uint32_t light_count;          /* reference lights + triangle lights */
    uint32_t tri_light_count;
    uint32_t node4_count;          /* 4-wide nodes (128 B each) of the same tree */
    uint32_t stack4;               /* traversal stack entries the 4-wide tree can need */
    uint64_t device_bytes;         /* nodes + triangles + shading records */
    float    build_ms;
    float    bounds_min[3], bounds_max[3];
    uint32_t bvh_depth4;           /* levels of the 4-wide tree (the stackless traversal takes up to 31) */
} rrh_scene_info;
int rrh_scene_upload(rrh_ctx* ctx, const rrh_vertex* verts, uint32_t nv, const uint32_t* idx,
                     uint32_t ni, const rrh_mesh* meshes, uint32_t nm, const rrh_material* mats,
                     uint32_t nmat, const rrh_instance* insts, uint32_t ninst,
                     const rrh_light* lights, uint32_t nl, uint32_t flags, rrh_scene** out);
int rrh_scene_destroy(rrh_scene* s);
int rrh_scene_info_get(const rrh_scene* s, rrh_scene_info* out);
/* world-space triangles (9 floats each) in flattening order, for checking */
int rrh_scene_triangles(const rrh_scene* s, float* out, uint32_t max_tris);
/* moves the instances of an uploaded scene (same count, same meshes; new transforms), then
 * refits or rebuilds its BVH on the device, updates the triangle-light records and bounds.
 * No reference counterpart: the reference builds its acceleration structures once per load
 * (Scene.cpp:449-507, AccelStructure.cpp:98-330); the refit plays the role of Vulkan's update
 * build mode.  Closest hits after either mode equal a fresh upload's bit for bit. */
#define RRH_UPDATE_REFIT 0         /* keep the tree, recompute every box bottom-up (quality drifts with motion) */
#define RRH_UPDATE_REBUILD 1       /* new tree by the context's builder (RRH_OPT_BVH_BUILDER) */
int rrh_scene_update_instances(rrh_ctx* ctx, rrh_scene* s, const rrh_instance* insts, uint32_t ninst, int mode);
/* replaces Scene::update_camera (src/scene/Scene.cpp:554-585) */
int rrh_camera_set(rrh_scene* s, const rrh_camera* cam);
/* host restatement of Camera::look_at + rebuild_matrices (src/scene/Camera.cpp:27-36,130-146) */
int rrh_camera_look_at(const float eye[3], const float center[3], float fov_y_deg, uint32_t width,
                       uint32_t height, float near_plane, float far_plane, uint32_t frame_index,
                       rrh_camera* out);

/* ---- procedural scenes (the reference ships only build_cornell_box,
 *      src/scene/Scene.cpp:144-251; the others are seeded generators) ----- */
#define RRH_GEN_CORNELL_REF 0      /* the reference Cornell box incl. its point light */
#define RRH_GEN_CORNELL_RANDOM 1   /* 36 tris, random axis-aligned blocks, albedo U[.1,.9], quad light L=15 */
#define RRH_GEN_HEIGHTFIELD 2      /* 512^2-cell fBm heightfield + 2000 icospheres + 4096 emissive tris */
#define RRH_GEN_INDOOR 3           /* 10x10x4 room, ~1M tris, 0.5x0.5 window, directional sun */
typedef struct {
    rrh_vertex* verts; uint32_t nv;
    uint32_t* idx; uint32_t ni;
    rrh_mesh* meshes; uint32_t nm;
    rrh_material* mats; uint32_t nmat;
    rrh_instance* insts; uint32_t ninst;
    rrh_light* lights; uint32_t nl;
    uint32_t flags;                /* suggested rrh_scene_upload flags */
    float cam_eye[3], cam_center[3], fov_y_deg;
} rrh_scene_desc;
/* scale: 1.0 = full size; smaller values shrink tessellation for tests */
int rrh_scene_generate(int kind, uint32_t seed, float scale, rrh_scene_desc* out);
int rrh_scene_desc_free(rrh_scene_desc* d);

/* ---- images: replace rr::rhi::Image + Readback (src/rhi/Readback.cpp:83-152) ---- */
#define RRH_FMT_RGBA32F 0
#define RRH_FMT_RGBA16F 1
#define RRH_FMT_RGBA64F 2          /* FP64 accumulation */
#define RRH_FMT_R32U 3
#define RRH_FMT_RGBA8 4            /* 8-bit RGBA (tonemapped display / screenshot) */
int rrh_image_create(rrh_ctx* ctx, uint32_t width, uint32_t height, int format, rrh_image** out);
int rrh_image_destroy(rrh_image* img);
int rrh_image_upload(rrh_ctx* ctx, rrh_image* img, const void* host, size_t bytes);
int rrh_image_download(rrh_ctx* ctx, const rrh_image* img, void* host, size_t bytes);
int rrh_image_clear(rrh_ctx* ctx, rrh_image* img);

/* ---- G-buffer: replaces GBufferPass::execute + gbuffer.slang (src/passes/gbuffer/
 *      GBufferPass.cpp:216-330): pixel-centre rays, back faces culled, depth within the
 *      camera's near / far planes; position RGBA32F (xyz, 1), normal RGBA32F (xyz,
 *      asfloat(material)), material id R32U; misses keep (0, 0, 0xFFFFFFFF) ------------- */
int rrh_gbuffer(rrh_ctx* ctx, rrh_scene* s, rrh_image* position, rrh_image* normal, rrh_image* material_id);
/* ---- screenshot tonemap: replaces the host loop of Application::capture_screenshot
 *      (src/app/Application.cpp:853-866): ACES (Narkowicz) of hdr * exposure, clamp, gamma
 *      1/2.2, 8 bits rounded to nearest (exactly); RGBA32F or RGBA64F in, RGBA8 out ------- */
int rrh_tonemap(rrh_ctx* ctx, const rrh_image* hdr, float exposure, rrh_image* ldr);
/* the 256 fp32 thresholds of that rounding (T_0 = 0; byte = largest k with T_k <= value), so
 * host code can produce the same bytes */
int rrh_tonemap_thresholds(float out[256]);

/* ---- closest hit / occlusion: the RayQuery TraceRayInline + Proceed loop
 *      (assets/shaders/passes/pathtracer_reference/pathtracer.slang:45-75) -- */
typedef struct { float origin[3]; float tmin; float dir[3]; float tmax; } rrh_ray;   /* 32 B */
typedef struct { float t; uint32_t prim; float b1, b2; } rrh_hit;  /* prim = flattened tri id, ~0u = miss */
int rrh_trace_closest(rrh_ctx* ctx, const rrh_scene* s, const rrh_ray* rays, uint32_t n, rrh_hit* hits);
int rrh_trace_occluded(rrh_ctx* ctx, const rrh_scene* s, const rrh_ray* rays, uint32_t n, uint32_t* occluded);
/* same on device-resident buffers (no copies), for timing */
int rrh_trace_closest_device(rrh_ctx* ctx, const rrh_scene* s, const void* d_rays, uint32_t n, void* d_hits);
/* device scratch for the _device variant: allocate / free n rays + hits */
int rrh_trace_buffers(rrh_ctx* ctx, uint32_t n, void** d_rays, void** d_hits);
int rrh_trace_buffers_free(rrh_ctx* ctx, void* d_rays, void* d_hits);
int rrh_copy_to_device(rrh_ctx* ctx, void* dst, const void* src, size_t bytes);
int rrh_copy_to_host(rrh_ctx* ctx, void* dst, const void* src, size_t bytes);

/* ---- path tracer: PathTracerPass::execute (src/passes/pathtracer_reference/
 *      PathTracerPass.cpp:170-224) + pathtracer.slang cs_main :181-265 ---- */
#define RRH_GUIDE_NONE 0
#define RRH_GUIDE_VMM 1            /* 8-lobe EM grid (rrh_grid of kind RRH_GRID_VMM) */
#define RRH_GUIDE_MLP 2            /* learned guide (rrh_mlp) */
typedef struct {
    uint32_t max_bounces;          /* PathTracerPass.h:46 default 5 */
    uint32_t frame_seed;           /* PathTracerPass.cpp:215 frame_count_++ */
    uint32_t rr_start;             /* Russian roulette from this bounce (pathtracer.slang:252 uses 3) */
    uint32_t guide_mode;           /* RRH_GUIDE_* */
    float    guide_alpha;          /* one-sample MIS probability of the guide */
    uint32_t record_vertices;      /* > 0: record path vertices, keep about this many as training samples */
    uint32_t flags;                /* RRH_PT_* */
    uint32_t guide_bounces;        /* guided modes: the guide samples bounces [0, guide_bounces), the
                                      BSDF alone after them; 0 = every bounce */
    uint32_t row_begin, row_end;   /* unguided frames: trace only the image rows [row_begin, row_end)
                                      (0, 0 = every row); the other rows of the radiance image are left
                                      as they are.  Every pixel's path depends only on its own seed, so
                                      bands rendered by separate calls (or contexts) equal one call
                                      over the whole frame bit for bit */
} rrh_pt_params;
#define RRH_PT_COUNT_FETCHES 1u    /* instrumented traversal: count node / triangle fetches (slower) */
typedef struct { uint64_t primary, extension, shadow; } rrh_ray_counts;
int rrh_pt_render(rrh_ctx* ctx, rrh_scene* s, const rrh_pt_params* p, rrh_image* radiance);
int rrh_pt_render_guided(rrh_ctx* ctx, rrh_scene* s, const rrh_pt_params* p, rrh_grid* grid,
                         rrh_mlp* mlp, rrh_image* radiance);
/* rrh_pt_render without any device allocation.  No reference counterpart: for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The ray queues, hits, shadow queue and counters live in one file-scope device
 * array per GPU (42 MB, held until the process ends) instead of context buffers; the result and
 * the ray counts equal rrh_pt_render's bit for bit.  Limits: at most 262,144 pixels (512x512;
 * larger extents return RRH_E_RANGE), unguided (guide_mode RRH_GUIDE_NONE, record_vertices 0),
 * flags 0, RRH_OPT_RAY_SORT off, no traversal profiling.  Every call on a GPU shares that GPU's
 * array (the last call's contents stay in it): calls that overlap on one GPU, from any context
 * or stream, race and must be ordered by the caller. */
int rrh_pt_render_noalloc(rrh_ctx* ctx, rrh_scene* s, const rrh_pt_params* p, rrh_image* radiance);
/* ray counts of the last render on this context (reads back device counters) */
int rrh_pt_ray_counts(rrh_ctx* ctx, rrh_ray_counts* out);
/* rays traced since the last reset by every render / ReSTIR DI call on this context,
 * summed on the device (no host round trip per frame); reset != 0 clears them */
int rrh_ray_totals(rrh_ctx* ctx, rrh_ray_counts* out, int reset);
/* node (64 B) and triangle (48 B) fetches of closest-hit traversal counted by renders
 * with RRH_PT_COUNT_FETCHES since the last reset, and the rays they cover */
int rrh_fetch_counts(rrh_ctx* ctx, uint64_t* node_fetches, uint64_t* tri_fetches, uint64_t* rays, int reset);
/* the same for the path tracer's shadow rays (occlusion traversal; counted for quantized 4-wide
 * trees with the speculative or the stackless traversal) */
int rrh_fetch_counts_shadow(rrh_ctx* ctx, uint64_t* node_fetches, uint64_t* tri_fetches, uint64_t* rays, int reset);
/* the same frames, per wave: the sum over waves of 64 x the largest per-ray node (and triangle)
 * count of the wave; node_fetches / closest_nodes bounds the lane utilisation the spread of ray
 * lengths allows (no reference counterpart; measurement only) */
int rrh_fetch_wave_max(rrh_ctx* ctx, uint64_t* closest_nodes, uint64_t* closest_tris, uint64_t* shadow_nodes,
                       uint64_t* shadow_tris, int reset);
/* closest-hit and shadow rays traced by rrh_pt_render* on this context since the last reset
 * (the path tracer's own rays: rrh_ray_totals also counts the ReSTIR DI pass) */
int rrh_pt_traversal_rays(rrh_ctx* ctx, uint64_t* closest, uint64_t* shadow, int reset);
/* GPU time of the path tracer's traversal launches (closest hit, paired closest hit + shadow,
 * shadow) inside rrh_pt_render, taken with HIP events on the context stream while profiling
 * is on */
/* tuning options of a context (no reference counterpart) */
#define RRH_OPT_TRAVERSAL_STACK 1   /* RRH_STACK_* : where the BVH traversal stack lives */
#define RRH_OPT_XCD_REMAP 2         /* 1: give each XCD runs of 16 consecutive workgroups (rotating) */
#define RRH_OPT_PERSISTENT 4        /* 1: the path tracer's ray queues traced by persistent waves that
                                       hand finished lanes new rays (binary nodes, and quantized 4-wide
                                       nodes with an LDS or private stack, where one launch drains a
                                       bounce's shadow and extension rays); 2..64: the same with at most
                                       that many waves per CU; 0 (default): one ray per lane */
#define RRH_OPT_REFILL 14           /* persistent 4-wide traversal: a wave takes new rays once this many
                                       of its 64 lanes are idle, 1..64 (default 16) */
#define RRH_OPT_RPT_STEP_VISITS 16  /* ReSTIR PT initial stage: 1..4096 node visits per traversal pass, a lane
                                       whose ray has finished advancing its candidate while the others'
                                       rays are in flight (speculative traversal); 0: every traversal
                                       call runs all the wave's rays to their end */
#define RRH_OPT_SUSPEND_VISITS 15   /* 1..4096: on the quantized 4-wide tree (speculative, LDS or private
                                       stack) the path tracer's rays stop after this many node visits
                                       and a second launch resumes them, compacted into full waves;
                                       0: every ray runs to its end in one launch */
#define RRH_OPT_SPECULATIVE 5       /* 1 (default): traversal parks leaves until the whole wave has one */
#define RRH_OPT_BVH_MAX_LEAF 6      /* largest BVH leaf, 1..8 (default 8); applies to later scene uploads */
#define RRH_OPT_BVH_TRAV_COST 7     /* SAH cost of a traversal step x 100 (default 100); later uploads */
#define RRH_OPT_BVH_NODES 3         /* RRH_NODES_*: the node format traversal reads (all share the leaves) */
#define RRH_OPT_BVH_BUILDER 8       /* RRH_BVH_BUILD_*: who builds the BVH of later scene uploads */
#define RRH_BVH_BUILD_HOST 0        /* binned SAH on the host threads (default) */
#define RRH_BVH_BUILD_DEVICE 1      /* PLOC clustering + SAH leaf collapse on the GPU (same formats, same hits) */
#define RRH_OPT_PLOC_RADIUS 9       /* device builder: clusters searched on each side, 1..32 (default 16) */
#define RRH_OPT_RAY_SORT 10         /* extension rays traced in the order of a key (direction octant, then
                                       origin Morton code); value = key bits sorted, 0 (off) or 4..32 */
#define RRH_OPT_RESTIR_CULL 11      /* ReSTIR PT / GI rays: 0 (default) two-sided, as the reference's instances
                                       (VK_GEOMETRY_INSTANCE_TRIANGLE_FACING_CULL_DISABLE_BIT_KHR,
                                       src/rhi/AccelStructure.cpp:90) override the shaders' cull flag;
                                       1: back faces culled, as RAY_FLAG_CULL_BACK_FACING_TRIANGLES alone */
#define RRH_OPT_TRAV_TOP_NODES 12   /* 0 (default) or 1..512: the path tracer's ray-queue kernels copy the
                                       first n quantized 4-wide nodes (the top of the tree: the nodes are
                                       numbered breadth-first there) into LDS per workgroup and read
                                       them from there */
#define RRH_OPT_TRACE_PAIRS 13      /* 1 (default): the path tracer traces each bounce's extension rays and the
                                       previous bounce's shadow rays in one launch (same results; not with
                                       RRH_PT_COUNT_FETCHES or RRH_OPT_TRAV_TOP_NODES); 0: one launch each */
#define RRH_NODES_BINARY 0          /* binary, 64 B (both child boxes in fp32) */
#define RRH_NODES_WIDE4 1           /* 4-wide, 128 B (fp32 child boxes) */
#define RRH_NODES_WIDE4Q 2          /* 4-wide, 64 B (8-bit child boxes, outward rounded); default */
#define RRH_STACK_PRIVATE 0         /* per-lane private (scratch) array, any depth */
#define RRH_STACK_LDS32 1           /* 32 entries per lane in LDS (when the tree can need more: hybrid16) */
#define RRH_STACK_HYBRID16 2        /* 16 entries in LDS + private overflow, any depth */
#define RRH_STACK_HYBRID32 3        /* 32 entries in LDS + private overflow, any depth */
#define RRH_STACK_NONE 4            /* stackless: the path tracer's ray queues are traced with 8 bits of
                                       trail per level in registers and parent links to climb back
                                       (4-wide trees up to 31 levels deep; other trees, and every
                                       other kernel, use the LDS stack RRH_STACK_LDS32 would give) */
int rrh_ctx_set_option(rrh_ctx* ctx, int option, int value);
/* plain device-to-device copy kernel (no reference counterpart): measures the achievable
 * HBM bandwidth on this device (2 x bytes / time, GB/s) and calibrates hardware
 * byte counters against a kernel of known traffic */
int rrh_stream_copy(rrh_ctx* ctx, size_t bytes, int iterations, float* gbps);
int rrh_profile_enable(rrh_ctx* ctx, int on);
int rrh_profile_traversal(rrh_ctx* ctx, float* total_ms, uint32_t* launches, int reset);
/* time origin for rrh_profile_spans: an event recorded on this context's stream (no reference
 * counterpart; measurement only) */
int rrh_profile_mark(rrh_ctx* ctx);
/* [begin, end] in ms after `origin`'s mark of each traversal launch profiled on `ctx` since the
 * last reset of rrh_profile_traversal (call before that reset); cap: slots in each array */
int rrh_profile_spans(rrh_ctx* ctx, const rrh_ctx* origin, float* begin_ms, float* end_ms, uint32_t cap,
                      uint32_t* n);

/* ---- accumulate: AccumulatePass::execute (src/passes/accumulate/
 *      AccumulatePass.cpp:167-218) + accumulate.slang:21-46 -------------- */
int rrh_accumulate(rrh_ctx* ctx, const rrh_image* radiance, rrh_image* accumulated,
                   uint32_t camera_moved, uint32_t accumulated_spp);
/* rrh_accumulate over image rows [row_begin, row_end) only (no reference counterpart: the row
 * band of a path-traced frame, rrh_pt_params.row_begin / row_end, accumulated on the stream of
 * the context that traced it, so no other stream has to wait for the band) */
int rrh_accumulate_rows(rrh_ctx* ctx, const rrh_image* radiance, rrh_image* accumulated,
                        uint32_t camera_moved, uint32_t accumulated_spp, uint32_t row_begin, uint32_t row_end);
/* FP64 running sum (RGBA64F); mean = sum / spp */
int rrh_accumulate_f64(rrh_ctx* ctx, const rrh_image* radiance, rrh_image* accum_sum,
                       uint32_t camera_moved, uint32_t accumulated_spp);

/* ---- ReSTIR DI: ReSTIRDIPass::execute (src/passes/restir_di/
 *      ReSTIRDIPass.cpp:306-447) + restir_di.slang cs_initial :186-262,
 *      cs_temporal :266-335, cs_spatial_shade :339-472 ------------------- */
typedef struct {
    uint32_t frame_seed;               /* frame_count_++ (ReSTIRDIPass.cpp:343) */
    uint32_t num_initial_candidates;   /* ReSTIRDIPass.h:91 default 16 */
    uint32_t temporal_enable;
    uint32_t spatial_enable;
    uint32_t spatial_num_neighbors;    /* default 4 */
    uint32_t spatial_radius;           /* pixels, default 15 */
    uint32_t _pad0, _pad1;
} rrh_di_params;
/* gbuf_pos/gbuf_norm: RGBA32F; reservoir: RGBA32F (light, w_sum, W, M) + its uv plane (R32U) */
int rrh_di_initial(rrh_ctx* ctx, rrh_scene* s, const rrh_di_params* p, rrh_image* gbuf_pos,
                   rrh_image* gbuf_norm, rrh_image* res, rrh_image* res_uv);
int rrh_di_temporal(rrh_ctx* ctx, rrh_scene* s, const rrh_di_params* p, const rrh_image* gbuf_pos,
                    const rrh_image* gbuf_norm, rrh_image* res_curr, rrh_image* res_curr_uv,
                    const rrh_image* res_prev, const rrh_image* res_prev_uv);
int rrh_di_spatial_shade(rrh_ctx* ctx, rrh_scene* s, const rrh_di_params* p,
                         const rrh_image* gbuf_pos, const rrh_image* gbuf_norm,
                         const rrh_image* res, const rrh_image* res_uv, rrh_image* out);
/* shadow rays traced by the last rrh_di_spatial_shade (one per shaded pixel) */
int rrh_di_ray_counts(rrh_ctx* ctx, rrh_ray_counts* out);

/* ---- ReSTIR PT, reconnection shift: ReSTIRPTPass::execute (src/passes/restir_pt/
 *      ReSTIRPTPass.cpp:529-813) + restir_pt.slang cs_initial :569-771, cs_temporal
 *      :775-872, cs_spatial_shade :876-1001 (guiding fit: rrh_grid_fit) ------------------
 * The G-buffer is ReSTIR DI's (position w < 0 on a miss, normal w = asfloat(material)).
 * A reservoir is four RGBA32F planes (restir_pt.slang:11-15):
 *   res[0] (reconnection position, w_sum)   res[1] (reconnection normal, asfloat(M))
 *   res[2] (suffix radiance, W)             res[3] (prefix throughput, valid)
 * and a history two (the frame's G-buffer position and normal, for reprojection).  Rays are
 * two-sided unless RRH_OPT_RESTIR_CULL is set.  grid: an RRH_GRID_HASH grid that guides the
 * first bounce and receives the splats, or NULL (cosine sampling, no splats). */
typedef struct {
    uint32_t frame_seed;               /* frame_count_++ (ReSTIRPTPass.cpp:561) */
    uint32_t num_initial_candidates;   /* ReSTIRPTPass.h:88 default 8 */
    uint32_t max_bounces;              /* path segments from the primary surface, default 4 */
    uint32_t temporal_enable;          /* 1 */
    uint32_t spatial_enable;           /* 1 */
    uint32_t spatial_num_neighbors;    /* 4 */
    uint32_t spatial_radius;           /* pixels, 12 */
    uint32_t include_direct_lighting;  /* 1: add the direct image (ReSTIR DI's output) */
    float    indirect_strength;        /* 1.0 */
    float    ray_bias;                 /* 1e-3 */
    uint32_t guiding_enable;           /* 1 */
    float    guiding_alpha;            /* one-sample MIS probability of the lobe, .5 */
} rrh_rpt_params;
/* cs_initial: candidates, resampling, the kept reservoir; writes the history planes */
int rrh_rpt_initial(rrh_ctx* ctx, rrh_scene* s, const rrh_rpt_params* p, rrh_grid* grid,
                    const rrh_image* gbuf_pos, const rrh_image* gbuf_norm, rrh_image* const res[4],
                    rrh_image* hist_pos, rrh_image* hist_norm);
/* cs_temporal: reuse of the previous frame's reservoir (prev_view_proj of the camera) */
int rrh_rpt_temporal(rrh_ctx* ctx, rrh_scene* s, const rrh_rpt_params* p, const rrh_image* gbuf_pos,
                     const rrh_image* gbuf_norm, rrh_image* const res[4], rrh_image* const res_prev[4],
                     const rrh_image* hist_prev_pos, const rrh_image* hist_prev_norm);
/* cs_spatial_shade: spatial reuse, shading into out (RGBA32F), splats into the grid;
 * direct: RGBA32F or NULL */
int rrh_rpt_spatial_shade(rrh_ctx* ctx, rrh_scene* s, const rrh_rpt_params* p, rrh_grid* grid,
                          const rrh_image* gbuf_pos, const rrh_image* gbuf_norm, const rrh_image* direct,
                          rrh_image* const res[4], rrh_image* out);
/* history reset (ReSTIRPTPass.cpp:616-663): reservoir planes zero, history position
 * (0, 0, 0, -1), history normal zero */
int rrh_rpt_clear(rrh_ctx* ctx, rrh_image* const res[4], rrh_image* hist_pos, rrh_image* hist_norm);
/* rays of the last frame: extension = the initial stage's path segments, shadow = its light
 * samples plus the temporal and spatial visibility rays */
int rrh_rpt_ray_counts(rrh_ctx* ctx, rrh_ray_counts* out);

/* ---- ReSTIR GI, one-bounce reconnection reuse: ReSTIRGIPass::execute (src/passes/restir_gi/
 *      ReSTIRGIPass.cpp:428-651) + restir_gi.slang cs_initial :566-684, cs_temporal :686-847,
 *      cs_spatial_shade :849-1013 ---------------------------------------------------------
 * G-buffer and history as ReSTIR PT's.  A reservoir is three RGBA32F planes
 * (restir_gi.slang:204-242): res[0] (sample position, w_sum), res[1] (sample radiance, W),
 * res[2] (sample normal, asfloat(M)); M = 0 is empty.  Cosine candidates, no guiding; rays
 * are two-sided unless RRH_OPT_RESTIR_CULL is set. */
typedef struct {
    uint32_t frame_seed;               /* frame_count_++ (ReSTIRGIPass.cpp:460) */
    uint32_t num_initial_candidates;   /* ReSTIRGIPass.h:72 default 8 */
    uint32_t temporal_enable;          /* 1 */
    uint32_t spatial_enable;           /* 1 */
    uint32_t spatial_num_neighbors;    /* 4 */
    uint32_t spatial_radius;           /* pixels, 12 */
    uint32_t include_direct_lighting;  /* 1 */
    float    indirect_strength;        /* 1.0 */
    float    ray_bias;                 /* 1e-3 */
    uint32_t _pad[3];
} rrh_rgi_params;
int rrh_rgi_initial(rrh_ctx* ctx, rrh_scene* s, const rrh_rgi_params* p, const rrh_image* gbuf_pos,
                    const rrh_image* gbuf_norm, rrh_image* const res[3], rrh_image* hist_pos, rrh_image* hist_norm);
int rrh_rgi_temporal(rrh_ctx* ctx, rrh_scene* s, const rrh_rgi_params* p, const rrh_image* gbuf_pos,
                     const rrh_image* gbuf_norm, rrh_image* const res[3], rrh_image* const res_prev[3],
                     const rrh_image* hist_prev_pos, const rrh_image* hist_prev_norm);
int rrh_rgi_spatial_shade(rrh_ctx* ctx, rrh_scene* s, const rrh_rgi_params* p, const rrh_image* gbuf_pos,
                          const rrh_image* gbuf_norm, const rrh_image* direct, rrh_image* const res[3],
                          rrh_image* out);
/* history reset (ReSTIRGIPass.cpp:380-400, 505-530): planes zero, history position (0, 0, 0, -1) */
int rrh_rgi_clear(rrh_ctx* ctx, rrh_image* const res[3], rrh_image* hist_pos, rrh_image* hist_norm);
/* rays of the last frame: extension = candidate segments, shadow = their light samples plus
 * the temporal and spatial visibility rays */
int rrh_rgi_ray_counts(rrh_ctx* ctx, rrh_ray_counts* out);

/* ---- A-trous: AtrousPass::execute (src/passes/denoise/AtrousPass.cpp:243-306)
 *      + atrous.slang cs_main :41-106 ------------------------------------- */
#define RRH_ATROUS_REFERENCE 0  /* RGBA32F radiance, RGBA32F position/normal G-buffer, Gaussian weights */
#define RRH_ATROUS_CONFIG 1     /* RGBA16F radiance+variance, RGBA16F normal+linear depth, SVGF weights */
typedef struct {
    uint32_t iterations;        /* clamped to [1, 7] (AtrousPass.cpp:269) */
    uint32_t mode;              /* RRH_ATROUS_* */
    float sigma_position;       /* reference: .25 ; config: sigma_depth 1.0 */
    float sigma_normal;         /* reference: .2  ; config: normal exponent 128 */
    float sigma_luminance;      /* reference: 1.0 ; config: sigma_color 4.0 (variance-scaled) */
    uint32_t flags;             /* RRH_ATROUS_FAST_MATH; 0 = exact */
    uint32_t _pad[2];
} rrh_atrous_params;
/* hardware exp / log / reciprocal / rsqrt in the per-tap weights instead of the correctly
 * rounded forms, within 1e-3 per pixel of the exact filter (no reference
 * counterpart; the reference's shader math is the GPU driver's approximate exp / divide) */
#define RRH_ATROUS_FAST_MATH 1u
/* runs `iterations` passes with stride 1<<i ping-ponging between tmp0/tmp1;
 * *out_slot receives the slot holding the result ((iterations-1)&1, AtrousPass.cpp:197-201).
 * Reference mode: gbuf_a = positions (w<0 miss), gbuf_b = normals.  Config mode: gbuf_a =
 * RGBA16F (n.xyz, linear depth), gbuf_b unused (may be NULL). */
int rrh_atrous(rrh_ctx* ctx, const rrh_atrous_params* p, const rrh_image* input,
               const rrh_image* gbuf_a, const rrh_image* gbuf_b, rrh_image* tmp0, rrh_image* tmp1,
               int* out_slot);
/* config-mode inputs from a DI frame: fp16 radiance + 3x3 luminance variance, fp16 normal + linear depth */
int rrh_atrous_prepare(rrh_ctx* ctx, const rrh_scene* s, const rrh_image* radiance,
                       const rrh_image* gbuf_pos, const rrh_image* gbuf_norm,
                       rrh_image* rad16, rrh_image* gbuf16);

/* ---- vMF guiding: the ReSTIR-PT guiding grid (assets/shaders/include/
 *      GuidingGrid.slang:54-189; restir_pt.slang:520-563,982-1015; host
 *      src/passes/restir_pt/ReSTIRPTPass.cpp:311-341,602-612,784-810) and the
 *      8-lobe EM dense grid ------------------------------------------------ */
#define RRH_GRID_HASH 0   /* reference: 2^20-cell hash grid, 1 lobe, int fixed-point splats */
#define RRH_GRID_VMM 1    /* res^3 dense grid, 8 lobes, FP64 sufficient statistics, online EM */
typedef struct {
    int kind;
    uint32_t cell_count;       /* HASH: 1<<20 (ReSTIRPTPass.h) */
    float base_cell_size;      /* HASH: .10 */
    float distance_slope;      /* HASH: .02 */
    uint32_t min_samples;      /* HASH: 8 */
    float kappa_max;           /* 512 */
    uint32_t res;              /* VMM: cells per axis (32) */
    uint32_t lobes;            /* VMM: 8 (max 8) */
    float bounds_min[3], bounds_max[3];   /* VMM grid box */
    float decay;               /* VMM: weight of previous statistics in online EM */
    float min_weight;          /* VMM: cells with less total weight keep their lobes */
} rrh_grid_params;
int rrh_grid_create(rrh_ctx* ctx, const rrh_grid_params* p, rrh_grid** out);
int rrh_grid_destroy(rrh_grid* g);
int rrh_grid_reset(rrh_ctx* ctx, rrh_grid* g);
/* HASH: guiding_cell_index + guiding_splat for n (position, bsdf id, direction) records */
int rrh_grid_splat(rrh_ctx* ctx, rrh_grid* g, const float* pos, const uint32_t* bsdf_ids,
                   const float* dirs, uint32_t n, const float cam_pos[3]);
/* HASH: cs_fit_guiding over all cells (fits and resets the accumulators) */
int rrh_grid_fit(rrh_ctx* ctx, rrh_grid* g);
/* HASH: sample_primary_dir for n vertices; seeds in/out */
int rrh_grid_sample_primary(rrh_ctx* ctx, rrh_grid* g, const float* pos, const float* nrm,
                            const uint32_t* bsdf_ids, uint32_t* seeds, uint32_t n,
                            const float cam_pos[3], float alpha, float* dirs, float* pdfs);
/* VMM: E-step over n weighted (pos, dir) samples + M-step; host arrays */
int rrh_grid_em_step(rrh_ctx* ctx, rrh_grid* g, const float* pos, const float* dirs,
                     const float* weights, uint32_t n);
/* VMM: E-step + M-step over the path vertices recorded by the last guided render */
int rrh_grid_em_from_vertices(rrh_ctx* ctx, rrh_grid* g);
/* read back state: HASH: fitted (4 floats/cell) and accum (4 u32/cell);
 * VMM: lobes (5 floats/lobe: mu3, kappa, pi) and stats (4 doubles/lobe) */
int rrh_grid_download(rrh_ctx* ctx, const rrh_grid* g, void* fitted_or_lobes, size_t bytes0,
                      void* accum_or_stats, size_t bytes1);
int rrh_grid_upload(rrh_ctx* ctx, rrh_grid* g, const void* fitted_or_lobes, size_t bytes0);

/* ---- learned guide (new; the reference guides with the single-lobe grid only).
 *      x(78) = 12-frequency sin/cos encoding of the normalised position (72) ++ n ++ wo;
 *      78 -> 64 -> 64 -> 64 -> 64 -> 40 ReLU MLP; 8 vMF lobes: mu = v/|v|,
 *      kappa = 0.1 + softplus(d) (<= 1000), pi = softmax(e).  BF16 weights and
 *      activations on MFMA with FP32 accumulation, FP32 master weights, Adam, dynamic
 *      loss scaling.  Parameters: W0 (64x78), b0, W1, b1, W2, b2, W3, b3, W4 (40x64), b4. */
#define RRH_MLP_PARAMS 20136
typedef struct {
    float bounds_min[3], bounds_max[3];   /* position normalisation box */
    float lr, beta1, beta2, eps;          /* Adam (1e-3, 0.9, 0.999, 1e-8) */
    uint32_t seed;                        /* weight initialisation */
    float loss_scale;                     /* initial loss scale (e.g. 1024) */
} rrh_mlp_params;
int rrh_mlp_create(rrh_ctx* ctx, const rrh_mlp_params* p, rrh_mlp** out);
int rrh_mlp_destroy(rrh_mlp* m);
int rrh_mlp_set_weights(rrh_ctx* ctx, rrh_mlp* m, const float* params, uint32_t n);
int rrh_mlp_get_weights(rrh_ctx* ctx, const rrh_mlp* m, float* params, uint32_t n);
/* lobes: n x 8 x (mu3, kappa, pi) */
int rrh_mlp_infer(rrh_ctx* ctx, rrh_mlp* m, const float* pos, const float* nrm, const float* wo, uint32_t n,
                  float* lobes);
/* mean weighted negative log-likelihood and its gradient (RRH_MLP_PARAMS floats), no update */
int rrh_mlp_loss_grad(rrh_ctx* ctx, rrh_mlp* m, const float* pos, const float* nrm, const float* wo,
                      const float* dirs, const float* w, uint32_t n, float* loss, float* grad);
/* one training step (forward, backward, Adam) on host samples / on the samples recorded by
 * the last rrh_pt_render_guided with record_vertices (at most `batch` of them) */
int rrh_mlp_train_step(rrh_ctx* ctx, rrh_mlp* m, const float* pos, const float* nrm, const float* wo,
                       const float* dirs, const float* w, uint32_t n, float* loss);
int rrh_mlp_train_from_vertices(rrh_ctx* ctx, rrh_mlp* m, uint32_t batch, float* loss);
/* one Adam step with a given (unscaled) gradient, for checking the optimiser */
int rrh_mlp_adam_apply(rrh_ctx* ctx, rrh_mlp* m, const float* grad, uint32_t n);
int rrh_mlp_state(rrh_ctx* ctx, const rrh_mlp* m, uint32_t* step, float* loss_scale);

#ifdef __cplusplus
}
#endif
#endif

// rrh_core.hip -- context, images, scene upload, camera and the ray-query entry points.
#include <chrono>
#include <cmath>
#include <cstring>

#include "rrh_internal.h"

namespace rrh {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

int DevBuf::ensure(size_t n) {
    if (n <= bytes && p) return RRH_OK;
    if (p) { (void)hipFree(p); p = nullptr; bytes = 0; }
    if (n == 0) n = 16;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) { p = nullptr; set_error(std::string("hipMalloc: ") + hipGetErrorString(e)); return RRH_E_NOMEM; }
    bytes = n;
    return RRH_OK;
}
void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

size_t format_bytes(int fmt) {
    switch (fmt) {
    case RRH_FMT_RGBA32F: return 16;
    case RRH_FMT_RGBA16F: return 8;
    case RRH_FMT_RGBA64F: return 32;
    case RRH_FMT_R32U: return 4;
    case RRH_FMT_RGBA8: return 4;
    default: return 0;
    }
}

}  // namespace rrh

using namespace rrh;

int rrh_scene::node_format() const { return ctx ? ctx->bvh_nodes : RRH_NODES_BINARY; }
uint32_t rrh_scene::stack_need() const { return node_format() == RRH_NODES_BINARY ? depth : stack4; }

rrh::DevScene rrh_scene::dev() const {
    DevScene d;
    d.nodes = (const float4*)nodes.p;
    d.nodes4 = (const float4*)nodes4.p;
    d.nodesq = (const uint4*)nodesq.p;
    d.parent4 = (const uint32_t*)parent4.p;
    d.node_fmt = (uint32_t)node_format();
    d.speculative = (ctx && ctx->speculative) ? 1u : 0u;
    d.tris = (const float4*)tris.p;
    d.tri_ref = (const uint2*)tri_ref.p;
    d.tri_vtx = (const uint4*)tri_vtx.p;
    d.tri_light = (const uint32_t*)tri_light.p;
    d.tl = (const float4*)tl.p;
    d.verts = (const rrh_vertex*)verts.p;
    d.idx = (const uint32_t*)idx.p;
    d.meshes = (const rrh_mesh*)meshes.p;
    d.mats = (const rrh_material*)mats.p;
    d.insts = (const rrh_instance*)insts.p;
    d.lights = (const rrh_light*)lights.p;
    d.cam = (const rrh_camera*)cam.p;
    d.nl = nl; d.ntl = ntl; d.ntri = ntri; d.flags = flags;
    return d;
}

template <class T>
static int upload_vec(DevBuf& b, const T* src, size_t n, hipStream_t st) {
    int rc = b.ensure(n * sizeof(T));
    if (rc) return rc;
    if (n) RRH_HIP(hipMemcpyAsync(b.p, src, n * sizeof(T), hipMemcpyHostToDevice, st));
    return RRH_OK;
}

extern "C" {

const char* rrh_last_error(void) { return g_err.c_str(); }

int rrh_ctx_create(int device, rrh_ctx** out) {
    RRH_ARG(out, "rrh_ctx_create: out is null");
    *out = nullptr;
    int n = 0;
    RRH_HIP(hipGetDeviceCount(&n));
    RRH_ARG(device >= 0 && device < n, "rrh_ctx_create: device index out of range");
    RRH_HIP(hipSetDevice(device));
    rrh_ctx* c = new rrh_ctx();
    c->device = device;
    {
        hipDeviceProp_t pr;
        if (hipGetDeviceProperties(&pr, device) == hipSuccess && pr.multiProcessorCount > 0) c->cu_count = pr.multiProcessorCount;
    }
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete c; set_error(hipGetErrorString(e)); return RRH_E_HIP; }
    for (int i = 0; i < 16; ++i) {
        e = hipEventCreate(&c->markers[i]);
        if (e != hipSuccess) { set_error(hipGetErrorString(e)); return RRH_E_HIP; }
    }
    e = hipEventCreateWithFlags(&c->join_event, hipEventDisableTiming);
    if (e != hipSuccess) { set_error(hipGetErrorString(e)); return RRH_E_HIP; }
    // device-side ray totals and fetch counters exist from the start, so no frame allocates them
    // (totals: primary, extension, shadow of every pass, then the path tracer's closest-hit and
    // shadow rays; fetch: closest-hit nodes, triangles, rays, then the same for shadow rays, then
    // the per-wave maxima of rrh_fetch_wave_max)
    int rc = c->totals.ensure(64);
    if (!rc) rc = c->fetch.ensure(128);
    if (!rc) rc = pt_scratch_init(device, &c->noalloc_scratch);
    if (rc) { rrh_ctx_destroy(c); return rc; }
    e = hipMemsetAsync(c->totals.p, 0, 64, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(c->fetch.p, 0, 128, c->stream);
    if (e != hipSuccess) { set_error(hipGetErrorString(e)); rrh_ctx_destroy(c); return RRH_E_HIP; }
    *out = c;
    return RRH_OK;
}

int rrh_ctx_destroy(rrh_ctx* c) {
    if (!c) return RRH_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (int i = 0; i < 2; ++i) { c->q_o[i].release(); c->q_d[i].release(); c->q_thp[i].release(); }
    c->hits.release(); c->sh_o.release(); c->sh_d.release(); c->sh_c.release();
    c->counters.release(); c->vtx.release(); c->di_counter.release(); c->di_pack.release(); c->rpt_counter.release(); c->rgi_counter.release(); c->scratch.release(); c->mlp_prep.release();
    c->train.release(); c->train_cnt.release(); c->fin_chunks.release(); c->fin_tmp.release(); c->totals.release(); c->fetch.release();
    c->sort_keys.release(); c->sort_perm.release(); c->sort_tmp.release(); c->susp.release();
    for (hipEvent_t e : c->prof_events) (void)hipEventDestroy(e);
    if (c->prof_mark) (void)hipEventDestroy(c->prof_mark);
    for (int i = 0; i < 16; ++i) (void)hipEventDestroy(c->markers[i]);
    if (c->join_event) (void)hipEventDestroy(c->join_event);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return RRH_OK;
}

int rrh_ctx_synchronize(rrh_ctx* c) {
    RRH_ARG(c, "null context");
    RRH_HIP(hipStreamSynchronize(c->stream));
    return RRH_OK;
}

int rrh_ctx_join(rrh_ctx* waiter, rrh_ctx* signaler) {
    RRH_ARG(waiter && signaler, "rrh_ctx_join: null context");
    RRH_ARG(waiter->device == signaler->device, "rrh_ctx_join: contexts on different devices");
    if (waiter == signaler) return RRH_OK;
    // the wait takes the event's state now; recording it again later does not affect it
    RRH_HIP(hipEventRecord(signaler->join_event, signaler->stream));
    RRH_HIP(hipStreamWaitEvent(waiter->stream, signaler->join_event, 0));
    return RRH_OK;
}

int rrh_device_info(rrh_ctx* c, int* cu_count, int* clock_khz, char* name, size_t name_len) {
    RRH_ARG(c, "null context");
    hipDeviceProp_t pr;
    RRH_HIP(hipGetDeviceProperties(&pr, c->device));
    if (cu_count) *cu_count = pr.multiProcessorCount;
    if (clock_khz) *clock_khz = pr.clockRate;
    if (name && name_len) { std::strncpy(name, pr.gcnArchName, name_len - 1); name[name_len - 1] = 0; }
    return RRH_OK;
}

int rrh_ctx_set_option(rrh_ctx* c, int option, int value) {
    RRH_ARG(c, "rrh_ctx_set_option: null context");
    if (option == RRH_OPT_TRAVERSAL_STACK) {
        RRH_ARG(value >= RRH_STACK_PRIVATE && value <= RRH_STACK_NONE, "rrh_ctx_set_option: bad stack kind");
        c->stack_kind = value;
        return RRH_OK;
    }
    if (option == RRH_OPT_TRAV_TOP_NODES) {
        RRH_ARG(value >= 0 && value <= 512, "rrh_ctx_set_option: top nodes out of 0..512");
        c->top_nodes = (uint32_t)value;
        return RRH_OK;
    }
    if (option == RRH_OPT_XCD_REMAP) { c->xcd_remap = value != 0; return RRH_OK; }
    if (option == RRH_OPT_RESTIR_CULL) { c->restir_cull = value != 0; return RRH_OK; }
    if (option == RRH_OPT_TRACE_PAIRS) { c->trace_pairs = value != 0; return RRH_OK; }
    if (option == RRH_OPT_PERSISTENT) {
        RRH_ARG(value >= 0 && value <= 64, "rrh_ctx_set_option: persistent waves per CU is 0 (off), 1 (all the CU holds) or 2..64");
        c->persistent = value != 0;
        c->persistent_waves = value > 1 ? (uint32_t)value : 1u;
        return RRH_OK;
    }
    if (option == RRH_OPT_SUSPEND_VISITS) {
        RRH_ARG(value >= 0 && value <= 4096, "rrh_ctx_set_option: suspend visits is 0 (off) or 1..4096 node visits");
        c->suspend_visits = (uint32_t)value;
        return RRH_OK;
    }
    if (option == RRH_OPT_RPT_STEP_VISITS) {
        RRH_ARG(value >= 0 && value <= 4096, "rrh_ctx_set_option: ReSTIR PT step visits is 0 (whole traversals) or 1..4096");
        c->rpt_step_visits = (uint32_t)value;
        return RRH_OK;
    }
    if (option == RRH_OPT_REFILL) {
        RRH_ARG(value >= 1 && value <= 64, "rrh_ctx_set_option: refill threshold is 1..64 idle lanes");
        c->refill = (uint32_t)value;
        return RRH_OK;
    }
    if (option == RRH_OPT_SPECULATIVE) { c->speculative = value != 0; return RRH_OK; }
    if (option == RRH_OPT_BVH_MAX_LEAF) {
        RRH_ARG(value >= 1 && value <= 8, "rrh_ctx_set_option: BVH leaf size is 1..8");
        c->bvh_max_leaf = (uint32_t)value;
        return RRH_OK;
    }
    if (option == RRH_OPT_BVH_TRAV_COST) {
        RRH_ARG(value >= 1 && value <= 10000, "rrh_ctx_set_option: traversal cost (x100) in [1, 10000]");
        c->bvh_trav_cost = (uint32_t)value;
        return RRH_OK;
    }
    if (option == RRH_OPT_BVH_BUILDER) {
        RRH_ARG(value == RRH_BVH_BUILD_HOST || value == RRH_BVH_BUILD_DEVICE, "rrh_ctx_set_option: bad BVH builder");
        c->bvh_builder = value;
        return RRH_OK;
    }
    if (option == RRH_OPT_RAY_SORT) {
        RRH_ARG(value == 0 || (value >= 4 && value <= 32), "rrh_ctx_set_option: ray sort bits are 0 or 4..32");
        c->ray_sort_bits = (uint32_t)value;
        return RRH_OK;
    }
    if (option == RRH_OPT_PLOC_RADIUS) {
        RRH_ARG(value >= 1 && value <= 32, "rrh_ctx_set_option: PLOC radius is 1..32");
        c->ploc_radius = (uint32_t)value;
        return RRH_OK;
    }
    if (option == RRH_OPT_BVH_NODES) {
        RRH_ARG(value >= RRH_NODES_BINARY && value <= RRH_NODES_WIDE4Q, "rrh_ctx_set_option: bad node format");
        c->bvh_nodes = value;
        return RRH_OK;
    }
    set_error("rrh_ctx_set_option: unknown option");
    return RRH_E_ARG;
}

int rrh_marker_record(rrh_ctx* c, int slot) {
    RRH_ARG(c && slot >= 0 && slot < 16, "rrh_marker_record: bad slot");
    RRH_HIP(hipEventRecord(c->markers[slot], c->stream));
    return RRH_OK;
}
int rrh_marker_elapsed(rrh_ctx* c, int a, int b, float* ms) {
    RRH_ARG(c && ms && a >= 0 && a < 16 && b >= 0 && b < 16, "rrh_marker_elapsed: bad args");
    RRH_HIP(hipEventSynchronize(c->markers[b]));
    RRH_HIP(hipEventElapsedTime(ms, c->markers[a], c->markers[b]));
    return RRH_OK;
}

// ---------------------------------------------------------------------------
// images
// ---------------------------------------------------------------------------
int rrh_image_create(rrh_ctx* c, uint32_t w, uint32_t h, int fmt, rrh_image** out) {
    RRH_ARG(c && out, "rrh_image_create: null argument");
    RRH_ARG(w > 0 && h > 0 && w <= 16384 && h <= 16384, "rrh_image_create: bad extent");
    size_t pb = format_bytes(fmt);
    RRH_ARG(pb > 0, "rrh_image_create: unknown format");
    rrh_image* im = new rrh_image();
    im->ctx = c; im->w = w; im->h = h; im->fmt = fmt;
    im->bytes = (size_t)w * h * pb;
    hipError_t e = hipMalloc(&im->p, im->bytes);
    if (e != hipSuccess) { delete im; set_error(std::string("hipMalloc: ") + hipGetErrorString(e)); return RRH_E_NOMEM; }
    e = hipMemsetAsync(im->p, 0, im->bytes, c->stream);
    if (e != hipSuccess) { (void)hipFree(im->p); delete im; set_error(hipGetErrorString(e)); return RRH_E_HIP; }
    *out = im;
    return RRH_OK;
}
int rrh_image_destroy(rrh_image* im) {
    if (!im) return RRH_OK;
    (void)hipDeviceSynchronize();  // the owning context may already be gone
    if (im->p) (void)hipFree(im->p);
    delete im;
    return RRH_OK;
}
int rrh_image_upload(rrh_ctx* c, rrh_image* im, const void* host, size_t bytes) {
    RRH_ARG(c && im && host, "rrh_image_upload: null argument");
    RRH_ARG(bytes == im->bytes, "rrh_image_upload: byte count does not match the image");
    RRH_HIP(hipMemcpyAsync(im->p, host, bytes, hipMemcpyHostToDevice, c->stream));
    RRH_HIP(hipStreamSynchronize(c->stream));
    return RRH_OK;
}
int rrh_image_download(rrh_ctx* c, const rrh_image* im, void* host, size_t bytes) {
    RRH_ARG(c && im && host, "rrh_image_download: null argument");
    RRH_ARG(bytes == im->bytes, "rrh_image_download: byte count does not match the image");
    RRH_HIP(hipMemcpyAsync(host, im->p, bytes, hipMemcpyDeviceToHost, c->stream));
    RRH_HIP(hipStreamSynchronize(c->stream));
    return RRH_OK;
}
int rrh_image_clear(rrh_ctx* c, rrh_image* im) {
    RRH_ARG(c && im, "rrh_image_clear: null argument");
    RRH_HIP(hipMemsetAsync(im->p, 0, im->bytes, c->stream));
    return RRH_OK;
}

int rrh_copy_to_device(rrh_ctx* c, void* dst, const void* src, size_t bytes) {
    RRH_ARG(c && dst && src, "rrh_copy_to_device: null argument");
    RRH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    RRH_HIP(hipStreamSynchronize(c->stream));
    return RRH_OK;
}
int rrh_copy_to_host(rrh_ctx* c, void* dst, const void* src, size_t bytes) {
    RRH_ARG(c && dst && src, "rrh_copy_to_host: null argument");
    RRH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    RRH_HIP(hipStreamSynchronize(c->stream));
    return RRH_OK;
}

// ---------------------------------------------------------------------------
// camera -- Camera.cpp:27-36,120-146 restated in double, GpuCameraData fill
// Scene.cpp:554-585 (glm column-major copied out row-major)
// ---------------------------------------------------------------------------
static void inv4(const double* m, double* out) {
    // Gauss-Jordan with partial pivoting
    double a[4][8];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 8; ++c) a[r][c] = (c < 4) ? m[r * 4 + c] : (c - 4 == r ? 1.0 : 0.0);
    for (int col = 0; col < 4; ++col) {
        int piv = col;
        for (int r = col + 1; r < 4; ++r)
            if (std::fabs(a[r][col]) > std::fabs(a[piv][col])) piv = r;
        for (int c = 0; c < 8; ++c) std::swap(a[col][c], a[piv][c]);
        double d = a[col][col];
        for (int c = 0; c < 8; ++c) a[col][c] /= d;
        for (int r = 0; r < 4; ++r) {
            if (r == col) continue;
            double f = a[r][col];
            for (int c = 0; c < 8; ++c) a[r][c] -= f * a[col][c];
        }
    }
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) out[r * 4 + c] = a[r][c + 4];
}

int rrh_camera_look_at(const float eye[3], const float center[3], float fov_y_deg, uint32_t width,
                       uint32_t height, float near_plane, float far_plane, uint32_t frame_index,
                       rrh_camera* out) {
    RRH_ARG(eye && center && out, "rrh_camera_look_at: null argument");
    RRH_ARG(width > 0 && height > 0, "rrh_camera_look_at: zero extent");
    RRH_ARG(fov_y_deg > 0.0f && fov_y_deg < 180.0f, "rrh_camera_look_at: fov out of range");
    const double PI = 3.14159265358979323846;
    double d[3];
    for (int i = 0; i < 3; ++i) d[i] = (double)center[i] - (double)eye[i];
    double dl = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    RRH_ARG(dl > 0.0, "rrh_camera_look_at: eye == center");
    for (double& x : d) x /= dl;
    double pitch = std::asin(std::max(-1.0, std::min(1.0, d[1])));
    double yaw = std::atan2(d[2], d[0]);
    double f[3] = {std::cos(pitch) * std::cos(yaw), std::sin(pitch), std::cos(pitch) * std::sin(yaw)};
    double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (double& x : f) x /= fl;
    double r[3] = {-f[2], 0.0, f[0]};  // cross(f, (0,1,0))
    double rl = std::sqrt(r[0] * r[0] + r[2] * r[2]);
    if (rl < 1e-5) { r[0] = 1; r[1] = 0; r[2] = 0; rl = 1; }
    for (double& x : r) x /= rl;
    double u[3] = {r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0]};
    double ul = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    for (double& x : u) x /= ul;
    // glm::lookAtRH: s = normalize(cross(f, up)), u' = cross(s, f)
    double s[3] = {f[1] * u[2] - f[2] * u[1], f[2] * u[0] - f[0] * u[2], f[0] * u[1] - f[1] * u[0]};
    double sl = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (double& x : s) x /= sl;
    double uu[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
    double e[3] = {eye[0], eye[1], eye[2]};
    double view[16] = {s[0], s[1], s[2], -(s[0] * e[0] + s[1] * e[1] + s[2] * e[2]),
                       uu[0], uu[1], uu[2], -(uu[0] * e[0] + uu[1] * e[1] + uu[2] * e[2]),
                       -f[0], -f[1], -f[2], (f[0] * e[0] + f[1] * e[1] + f[2] * e[2]),
                       0, 0, 0, 1};
    double aspect = (double)width / (double)height;
    double th = std::tan((double)fov_y_deg * PI / 180.0 * 0.5);
    double n = near_plane, fr = far_plane;
    double proj[16] = {0};
    proj[0] = 1.0 / (aspect * th);
    proj[5] = -1.0 / th;  // Vulkan Y flip (Camera.cpp:142)
    proj[10] = -(fr + n) / (fr - n);
    proj[11] = -(2.0 * fr * n) / (fr - n);
    proj[14] = -1.0;
    double iv[16], ip[16], vp[16];
    inv4(view, iv);
    inv4(proj, ip);
    for (int rr = 0; rr < 4; ++rr)
        for (int cc = 0; cc < 4; ++cc) {
            double acc = 0;
            for (int k = 0; k < 4; ++k) acc += proj[rr * 4 + k] * view[k * 4 + cc];
            vp[rr * 4 + cc] = acc;
        }
    std::memset(out, 0, sizeof(*out));
    for (int i = 0; i < 16; ++i) {
        out->view[i] = (float)view[i];
        out->inv_view[i] = (float)iv[i];
        out->inv_proj[i] = (float)ip[i];
        out->prev_view_proj[i] = (float)vp[i];
    }
    for (int i = 0; i < 3; ++i) { out->position[i] = eye[i]; out->direction[i] = (float)f[i]; }
    out->near_plane = near_plane;
    out->far_plane = far_plane;
    out->frame_index = frame_index;
    out->screen_width = width;
    out->screen_height = height;
    return RRH_OK;
}

// ---------------------------------------------------------------------------
// scene upload
// ---------------------------------------------------------------------------
static inline float mrow_h(const float* M, int r, float x, float y, float z, float w) {
    return M[r * 4 + 0] * x + M[r * 4 + 1] * y + M[r * 4 + 2] * z + M[r * 4 + 3] * w;
}


// host binned-SAH build of s->tri_world, uploaded into the scene's node and triangle buffers
static int host_build_upload(rrh_ctx* c, rrh_scene* s) {
    BvhOut bvh;
    std::string err;
    if (!build_bvh(s->tri_world.data(), s->ntri, bvh, err, c->bvh_max_leaf, (float)c->bvh_trav_cost * 0.01f)) {
        set_error("rrh_scene_upload: " + err);
        return RRH_E_STATE;
    }
    s->nnodes = (uint32_t)(bvh.nodes.size() / 4);
    s->depth = bvh.depth;
    s->leaves = bvh.leaves;
    s->nnodes4 = (uint32_t)(bvh.nodes4.size() / 8);
    s->depth4 = bvh.depth4;
    s->stack4 = bvh.stack4;
    s->refit_ready = false;
    hipStream_t st = c->stream;
    int rc = RRH_OK;
    if (!rc) rc = upload_vec(s->nodes, bvh.nodes.data(), bvh.nodes.size(), st);
    if (!rc) rc = upload_vec(s->nodes4, bvh.nodes4.data(), bvh.nodes4.size(), st);
    if (!rc) rc = upload_vec(s->nodesq, bvh.nodesq.data(), bvh.nodesq.size(), st);
    if (!rc) rc = upload_vec(s->tris, bvh.tris.data(), bvh.tris.size(), st);
    if (!rc) rc = tree_links(c, s, kTopBfsNodes);
    if (!rc) {
        hipError_t e = hipStreamSynchronize(st);  // the host vectors die with this frame
        if (e != hipSuccess) { set_error(hipGetErrorString(e)); rc = RRH_E_HIP; }
    }
    return rc;
}

int rrh_scene_upload(rrh_ctx* c, const rrh_vertex* verts, uint32_t nv, const uint32_t* idx, uint32_t ni,
                     const rrh_mesh* meshes, uint32_t nm, const rrh_material* mats, uint32_t nmat,
                     const rrh_instance* insts, uint32_t ninst, const rrh_light* lights, uint32_t nl,
                     uint32_t flags, rrh_scene** out) {
    RRH_ARG(c && out, "rrh_scene_upload: null argument");
    RRH_ARG(verts && idx && meshes && mats && insts, "rrh_scene_upload: null scene array");
    RRH_ARG(nl == 0 || lights, "rrh_scene_upload: lights is null");
    *out = nullptr;
    // validate references (the reference trusts its own data; a C ABI cannot)
    for (uint32_t m = 0; m < nm; ++m) {
        const rrh_mesh& me = meshes[m];
        RRH_ARG(me.material_index < nmat, "rrh_scene_upload: mesh material index out of range");
        RRH_ARG(me.index_count % 3 == 0, "rrh_scene_upload: mesh index_count not a multiple of 3");
        RRH_ARG(me.index_byte_offset % 4 == 0, "rrh_scene_upload: misaligned index_byte_offset");
        RRH_ARG((uint64_t)me.index_byte_offset / 4 + me.index_count <= ni, "rrh_scene_upload: mesh indices out of range");
        for (uint32_t k = 0; k < me.index_count; ++k) {
            uint64_t vi = (uint64_t)me.vertex_base + idx[me.index_byte_offset / 4 + k];
            RRH_ARG(vi < nv, "rrh_scene_upload: vertex index out of range");
        }
    }
    uint64_t ntri64 = 0;
    for (uint32_t i = 0; i < ninst; ++i) {
        RRH_ARG(insts[i].mesh_index < nm, "rrh_scene_upload: instance mesh index out of range");
        ntri64 += meshes[insts[i].mesh_index].index_count / 3;
    }
    RRH_ARG(ntri64 > 0, "rrh_scene_upload: scene has no triangles");
    if (ntri64 >= (1ull << 28)) { set_error("rrh_scene_upload: more than 2^28 triangles"); return RRH_E_RANGE; }
    uint32_t ntri = (uint32_t)ntri64;
    auto t0 = std::chrono::steady_clock::now();
    rrh_scene* s = new rrh_scene();
    s->ctx = c;
    s->flags = flags;
    s->ntri = ntri;
    s->nl = nl;
    s->tri_world.resize((size_t)ntri * 9);
    std::vector<uint2> tri_ref(ntri);
    std::vector<uint4> tri_vtx(ntri);
    std::vector<uint32_t> tri_light(ntri, ~0u);
    std::vector<float4> tl;
    uint32_t t = 0;
    for (uint32_t i = 0; i < ninst; ++i) {
        const rrh_mesh& me = meshes[insts[i].mesh_index];
        const float* T = insts[i].transform;
        const rrh_material& mt = mats[me.material_index];
        bool em = mt.emissive[0] > 0.0f || mt.emissive[1] > 0.0f || mt.emissive[2] > 0.0f;
        for (uint32_t p = 0; p < me.index_count / 3; ++p, ++t) {
            uint32_t base = me.index_byte_offset / 4 + p * 3;
            float* w = &s->tri_world[(size_t)t * 9];
            for (int k = 0; k < 3; ++k) {
                const rrh_vertex& vv = verts[me.vertex_base + idx[base + k]];
                w[k * 3 + 0] = mrow_h(T, 0, vv.pos[0], vv.pos[1], vv.pos[2], 1.0f);
                w[k * 3 + 1] = mrow_h(T, 1, vv.pos[0], vv.pos[1], vv.pos[2], 1.0f);
                w[k * 3 + 2] = mrow_h(T, 2, vv.pos[0], vv.pos[1], vv.pos[2], 1.0f);
            }
            tri_ref[t] = make_uint2(i, p);
            tri_vtx[t] = make_uint4(me.vertex_base + idx[base], me.vertex_base + idx[base + 1],
                                    me.vertex_base + idx[base + 2], i);
            if ((flags & RRH_SCENE_TRI_LIGHTS) && em) {
                tri_light[t] = nl + (uint32_t)(tl.size() / kTlF4);
                float4 a = {w[0], w[1], w[2], mt.emissive[0]};
                float4 b = {w[3], w[4], w[5], mt.emissive[1]};
                float4 cc = {w[6], w[7], w[8], mt.emissive[2]};
                // e1 x e2, its length, area and unit normal: the kernels' fp32 expressions
                // (cross, length = sqrtf(x*x + y*y + z*z), scale by 1 / length), same order
                const float e1x = w[3] - w[0], e1y = w[4] - w[1], e1z = w[5] - w[2];
                const float e2x = w[6] - w[0], e2y = w[7] - w[1], e2z = w[8] - w[2];
                const float crx = e1y * e2z - e1z * e2y, cry = e1z * e2x - e1x * e2z, crz = e1x * e2y - e1y * e2x;
                const float l2 = sqrtf(crx * crx + cry * cry + crz * crz);
                const float inv = 1.0f / l2;
                float4 d = {0.0f, 0.5f * l2, 0.0f, 0.0f};
                std::memcpy(&d.x, &t, 4);
                float4 nrm = {crx * inv, cry * inv, crz * inv, 0.0f};
                tl.push_back(a); tl.push_back(b); tl.push_back(cc); tl.push_back(d); tl.push_back(nrm);
            }
        }
    }
    s->ntl = (uint32_t)(tl.size() / kTlF4);
    for (int k = 0; k < 3; ++k) { s->bmin[k] = INFINITY; s->bmax[k] = -INFINITY; }
    for (size_t i = 0; i < (size_t)ntri * 3; ++i)
        for (int k = 0; k < 3; ++k) {
            s->bmin[k] = std::min(s->bmin[k], s->tri_world[i * 3 + k]);
            s->bmax[k] = std::max(s->bmax[k], s->tri_world[i * 3 + k]);
        }
    hipStream_t st = c->stream;
    int rc = RRH_OK;
    s->ninst = ninst;
    s->inst_mesh.resize(ninst);
    for (uint32_t i = 0; i < ninst; ++i) s->inst_mesh[i] = insts[i].mesh_index;
    if (c->bvh_builder == RRH_BVH_BUILD_DEVICE) {
        rc = upload_vec(s->twd, s->tri_world.data(), s->tri_world.size(), st);
        if (!rc) rc = build_bvh_device(c, s, (const float*)s->twd.p, c->bvh_max_leaf, (float)c->bvh_trav_cost * 0.01f,
                                       c->ploc_radius);
    } else {
        rc = host_build_upload(c, s);
    }
    if (rc) {
        const std::string msg = std::string(rrh_last_error()).rfind("rrh_scene_upload", 0) == 0
                                    ? std::string(rrh_last_error())
                                    : std::string("rrh_scene_upload: ") + rrh_last_error();
        rrh_scene_destroy(s);
        set_error(msg);
        return rc;
    }
    if (!rc) rc = upload_vec(s->tri_ref, tri_ref.data(), tri_ref.size(), st);
    if (!rc) rc = upload_vec(s->tri_vtx, tri_vtx.data(), tri_vtx.size(), st);
    if (!rc) rc = upload_vec(s->tri_light, tri_light.data(), tri_light.size(), st);
    if (!rc) rc = upload_vec(s->tl, tl.data(), tl.size(), st);
    if (!rc) rc = upload_vec(s->verts, verts, nv, st);
    if (!rc) rc = upload_vec(s->idx, idx, ni, st);
    if (!rc) rc = upload_vec(s->meshes, meshes, nm, st);
    if (!rc) rc = upload_vec(s->mats, mats, nmat, st);
    if (!rc) rc = upload_vec(s->insts, insts, ninst, st);
    if (!rc) rc = upload_vec(s->lights, lights, nl, st);
    if (!rc) rc = s->cam.ensure(sizeof(rrh_camera));
    if (!rc) {
        hipError_t e = hipStreamSynchronize(st);
        if (e != hipSuccess) { set_error(hipGetErrorString(e)); rc = RRH_E_HIP; }
    }
    if (rc) { rrh_scene_destroy(s); return rc; }
    s->build_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *out = s;
    return RRH_OK;
}

int rrh_scene_destroy(rrh_scene* s) {
    if (!s) return RRH_OK;
    (void)hipDeviceSynchronize();  // the owning context may already be gone
    DevBuf* bufs[] = {&s->nodes, &s->nodes4, &s->nodesq, &s->tris, &s->tri_ref, &s->tri_light, &s->tl, &s->verts, &s->idx,
                      &s->meshes, &s->mats, &s->insts, &s->lights, &s->cam, &s->twd, &s->refit, &s->tri_vtx};
    for (DevBuf* b : bufs) b->release();
    delete s;
    return RRH_OK;
}

int rrh_scene_info_get(const rrh_scene* s, rrh_scene_info* o) {
    RRH_ARG(s && o, "rrh_scene_info_get: null argument");
    std::memset(o, 0, sizeof(*o));
    o->tri_count = s->ntri;
    o->node_count = s->nnodes;
    o->bvh_depth = s->depth;
    o->leaf_count = s->leaves;
    o->light_count = s->nl + s->ntl;
    o->tri_light_count = s->ntl;
    o->node4_count = s->nnodes4;
    o->stack4 = s->stack4;
    o->device_bytes = (uint64_t)s->nnodes * 64 + (uint64_t)s->nnodes4 * 192 + (uint64_t)s->ntri * 48 +
                      (uint64_t)s->ntri * 12 + (uint64_t)s->ntl * kTlF4 * 16;
    o->build_ms = s->build_ms;
    for (int k = 0; k < 3; ++k) { o->bounds_min[k] = s->bmin[k]; o->bounds_max[k] = s->bmax[k]; }
    o->bvh_depth4 = s->depth4;
    return RRH_OK;
}

// the host copy of the world triangles, refreshed from the device after an instance update
static int sync_tri_world(const rrh_scene* s) {
    if (!s->tri_world_stale) return RRH_OK;
    RRH_HIP(hipMemcpyAsync(s->tri_world.data(), s->twd.p, (size_t)s->ntri * 36, hipMemcpyDeviceToHost, s->ctx->stream));
    RRH_HIP(hipStreamSynchronize(s->ctx->stream));
    s->tri_world_stale = false;
    return RRH_OK;
}

int rrh_scene_triangles(const rrh_scene* s, float* out, uint32_t max_tris) {
    RRH_ARG(s && out, "rrh_scene_triangles: null argument");
    RRH_ARG(max_tris >= s->ntri, "rrh_scene_triangles: output too small");
    int rc = sync_tri_world(s);
    if (rc) return rc;
    std::memcpy(out, s->tri_world.data(), (size_t)s->ntri * 36);
    return RRH_OK;
}

int rrh_scene_update_instances(rrh_ctx* c, rrh_scene* s, const rrh_instance* insts, uint32_t ninst, int mode) {
    RRH_ARG(c && s && insts, "rrh_scene_update_instances: null argument");
    RRH_ARG(s->ctx == c, "rrh_scene_update_instances: scene belongs to another context");
    RRH_ARG(mode == RRH_UPDATE_REFIT || mode == RRH_UPDATE_REBUILD, "rrh_scene_update_instances: bad mode");
    RRH_ARG(ninst == s->ninst, "rrh_scene_update_instances: instance count differs from the upload");
    for (uint32_t i = 0; i < ninst; ++i)
        RRH_ARG(insts[i].mesh_index == s->inst_mesh[i], "rrh_scene_update_instances: an instance changed its mesh");
    auto t0 = std::chrono::steady_clock::now();
    hipStream_t st = c->stream;
    RRH_HIP(hipMemcpyAsync(s->insts.p, insts, (size_t)ninst * sizeof(rrh_instance), hipMemcpyHostToDevice, st));
    int rc = flatten_device(c, s);
    if (rc) return rc;
    if (mode == RRH_UPDATE_REFIT) {
        rc = refit_bvh_device(c, s);
    } else if (c->bvh_builder == RRH_BVH_BUILD_DEVICE) {
        rc = build_bvh_device(c, s, (const float*)s->twd.p, c->bvh_max_leaf, (float)c->bvh_trav_cost * 0.01f,
                              c->ploc_radius);
    } else {
        rc = sync_tri_world(s);
        if (!rc) rc = host_build_upload(c, s);
    }
    if (rc) return rc;
    RRH_HIP(hipStreamSynchronize(st));
    s->build_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return RRH_OK;
}

int rrh_camera_set(rrh_scene* s, const rrh_camera* cam) {
    RRH_ARG(s && cam, "rrh_camera_set: null argument");
    RRH_ARG(cam->screen_width > 0 && cam->screen_height > 0, "rrh_camera_set: zero extent");
    s->cam_host = *cam;
    s->cam_set = true;
    RRH_HIP(hipMemcpyAsync(s->cam.p, cam, sizeof(rrh_camera), hipMemcpyHostToDevice, s->ctx->stream));
    return RRH_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// ray-query kernels (parity probe and timing)
// ---------------------------------------------------------------------------
// batch ray queries: private stack (STK = RRH_STACK_PRIVATE) or stackless (RRH_STACK_NONE,
// 4-wide trees), TOP: the first top_n quantized nodes read from LDS (dynamic shared memory)
extern __shared__ uint4 s_top_nodes[];
template <int STK, bool TOP>
__global__ void __launch_bounds__(256) k_trace_closest(DevScene S, const float4* __restrict__ rays, uint32_t n,
                                                      float4* __restrict__ hits, uint32_t top_n) {
    const auto top = stage_top<TOP>(S, s_top_nodes, top_n);
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 a = rays[2 * i], b = rays[2 * i + 1];
    typename StackSel<STK>::T st;
    Hit h = trace_closest_s<false>(S, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), a.w, b.w, st, nullptr, nullptr, top);
    hits[i] = make_float4(h.t, __uint_as_float(h.prim), h.b1, h.b2);
}
template <int STK, bool TOP>
__global__ void __launch_bounds__(256) k_trace_occluded(DevScene S, const float4* __restrict__ rays, uint32_t n,
                                                       uint32_t* __restrict__ occ, uint32_t top_n) {
    const auto top = stage_top<TOP>(S, s_top_nodes, top_n);
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 a = rays[2 * i], b = rays[2 * i + 1];
    typename StackSel<STK>::T st;
    occ[i] = trace_occluded_s(S, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), a.w, b.w, st, top) ? 1u : 0u;
}
// launches a batch query with the context's traversal options: stackless when RRH_STACK_NONE
// is set and the tree is 4-wide and at most 31 levels deep, the top of the tree in LDS when
// RRH_OPT_TRAV_TOP_NODES is set on quantized nodes
template <bool OCC>
hipError_t launch_query(const rrh_ctx* c, const rrh_scene* s, const void* rays, uint32_t n, void* out) {
    const int fmt = s->node_format();
    const bool sl = c->stack_kind == RRH_STACK_NONE && fmt != RRH_NODES_BINARY && s->depth4 <= 31;
    const uint32_t top_n = fmt == RRH_NODES_WIDE4Q ? std::min(c->top_nodes, s->top4) : 0u;
    const size_t lds = (size_t)top_n * 4 * sizeof(uint4);
    const dim3 g((n + 255) / 256), b(256);
    const DevScene S = s->dev();
    const float4* r = (const float4*)rays;
    auto go = [&](auto kc, auto ko) {
        if constexpr (OCC) hipLaunchKernelGGL(ko, g, b, lds, c->stream, S, r, n, (uint32_t*)out, top_n);
        else hipLaunchKernelGGL(kc, g, b, lds, c->stream, S, r, n, (float4*)out, top_n);
    };
    if (sl && top_n) go(k_trace_closest<RRH_STACK_NONE, true>, k_trace_occluded<RRH_STACK_NONE, true>);
    else if (sl) go(k_trace_closest<RRH_STACK_NONE, false>, k_trace_occluded<RRH_STACK_NONE, false>);
    else if (top_n) go(k_trace_closest<RRH_STACK_PRIVATE, true>, k_trace_occluded<RRH_STACK_PRIVATE, true>);
    else go(k_trace_closest<RRH_STACK_PRIVATE, false>, k_trace_occluded<RRH_STACK_PRIVATE, false>);
    return hipGetLastError();
}

// grid-stride copy, four independent 16 B nontemporal loads in flight per lane before their
// stores (measured: 5.28 TB/s on 32768 blocks against 4.56 TB/s for one load per iteration)
typedef float f4v __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) k_stream_copy(const f4v* __restrict__ src, f4v* __restrict__ dst, size_t n) {
    const size_t st = (size_t)gridDim.x * blockDim.x;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * st < n; i += 4 * st) {
        const f4v a = __builtin_nontemporal_load(src + i), b = __builtin_nontemporal_load(src + i + st);
        const f4v c = __builtin_nontemporal_load(src + i + 2 * st), d = __builtin_nontemporal_load(src + i + 3 * st);
        __builtin_nontemporal_store(a, dst + i);
        __builtin_nontemporal_store(b, dst + i + st);
        __builtin_nontemporal_store(c, dst + i + 2 * st);
        __builtin_nontemporal_store(d, dst + i + 3 * st);
    }
    for (; i < n; i += st) dst[i] = src[i];
}

extern "C" {

int rrh_stream_copy(rrh_ctx* c, size_t bytes, int iterations, float* gbps) {
    RRH_ARG(c && gbps && iterations > 0, "rrh_stream_copy: bad argument");
    RRH_ARG(bytes >= 4096 && bytes % 16 == 0, "rrh_stream_copy: bytes must be a multiple of 16, >= 4096");
    void *a = nullptr, *b = nullptr;
    RRH_HIP(hipMalloc(&a, bytes));
    hipError_t e = hipMalloc(&b, bytes);
    if (e != hipSuccess) { (void)hipFree(a); set_error(hipGetErrorString(e)); return RRH_E_NOMEM; }
    size_t n = bytes / 16;
    int grid = 32768;
    (void)hipMemsetAsync(a, 0, bytes, c->stream);
    hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, c->stream, (const f4v*)a, (f4v*)b, n);  // warm-up
    (void)hipEventRecord(c->markers[14], c->stream);
    for (int k = 0; k < iterations; ++k)
        hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, c->stream, (const f4v*)a, (f4v*)b, n);
    (void)hipEventRecord(c->markers[15], c->stream);
    e = hipEventSynchronize(c->markers[15]);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->markers[14], c->markers[15]);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) { set_error(hipGetErrorString(e)); return RRH_E_HIP; }
    *gbps = (float)(2.0 * (double)bytes * iterations / (ms * 1e-3) / 1e9);
    return RRH_OK;
}

int rrh_trace_closest_device(rrh_ctx* c, const rrh_scene* s, const void* d_rays, uint32_t n, void* d_hits) {
    RRH_ARG(c && s && d_rays && d_hits, "rrh_trace_closest_device: null argument");
    if (n == 0) return RRH_OK;
    RRH_HIP(launch_query<false>(c, s, d_rays, n, d_hits));
    return RRH_OK;
}

int rrh_trace_buffers(rrh_ctx* c, uint32_t n, void** d_rays, void** d_hits) {
    RRH_ARG(c && d_rays && d_hits && n > 0, "rrh_trace_buffers: bad argument");
    RRH_HIP(hipMalloc(d_rays, (size_t)n * 32));
    hipError_t e = hipMalloc(d_hits, (size_t)n * 16);
    if (e != hipSuccess) { (void)hipFree(*d_rays); set_error(hipGetErrorString(e)); return RRH_E_NOMEM; }
    return RRH_OK;
}
int rrh_trace_buffers_free(rrh_ctx* c, void* d_rays, void* d_hits) {
    RRH_ARG(c, "null context");
    (void)hipStreamSynchronize(c->stream);
    if (d_rays) (void)hipFree(d_rays);
    if (d_hits) (void)hipFree(d_hits);
    return RRH_OK;
}

int rrh_trace_closest(rrh_ctx* c, const rrh_scene* s, const rrh_ray* rays, uint32_t n, rrh_hit* hits) {
    RRH_ARG(c && s && rays && hits, "rrh_trace_closest: null argument");
    if (n == 0) return RRH_OK;
    void *dr = nullptr, *dh = nullptr;
    int rc = rrh_trace_buffers(c, n, &dr, &dh);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(dr, rays, (size_t)n * 32, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        e = launch_query<false>(c, s, dr, n, dh);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(hits, dh, (size_t)n * 16, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dr);
    (void)hipFree(dh);
    if (e != hipSuccess) { set_error(std::string("rrh_trace_closest: ") + hipGetErrorString(e)); return RRH_E_HIP; }
    return RRH_OK;
}

int rrh_trace_occluded(rrh_ctx* c, const rrh_scene* s, const rrh_ray* rays, uint32_t n, uint32_t* occ) {
    RRH_ARG(c && s && rays && occ, "rrh_trace_occluded: null argument");
    if (n == 0) return RRH_OK;
    void *dr = nullptr, *dh = nullptr;
    int rc = rrh_trace_buffers(c, n, &dr, &dh);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(dr, rays, (size_t)n * 32, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        e = launch_query<true>(c, s, dr, n, dh);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(occ, dh, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dr);
    (void)hipFree(dh);
    if (e != hipSuccess) { set_error(std::string("rrh_trace_occluded: ") + hipGetErrorString(e)); return RRH_E_HIP; }
    return RRH_OK;
}

}  // extern "C"

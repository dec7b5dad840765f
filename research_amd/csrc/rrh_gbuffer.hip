// rrh_gbuffer.hip -- the G-buffer and the screenshot tonemap on the device.
//
// G-buffer: replaces the reference's raster GBufferPass (GBufferPass.cpp:216-330,
// gbuffer.slang).  CDNA4 has no rasterizer; one ray through each pixel centre against the
// BVH gives the same coverage (pixel-centre sampling), the same visibility (nearest surface
// between the near and far planes, back faces culled like the pipeline's CullMode::Back with
// counter-clockwise front faces) and the same attributes: the world position blended from the
// transformed vertices (w = 1), the normal blended from the per-vertex normalize(M3 n) and
// normalized again (w = asfloat(material)), the material id; a miss keeps the clear values
// (0, 0, 0xFFFFFFFF).  One thread per pixel, quantized 4-wide nodes, private stack.
//
// Tonemap: replaces the host loop of the screenshot path (Application.cpp:853-866): ACES
// (Narkowicz) on the exposed value, clamp to [0, 1], then v^(1/2.2) * 255 rounded to nearest.
// The last step is taken as the count of thresholds T_k <= v, T_k = ((k - 1/2) / 255)^2.2
// rounded up to fp32 (computed once in double), which is the exact rounding of the formula
// and the same bytes on the host (rrh_tonemap_thresholds feeds the host layer's copy).
#include <cmath>
#include <mutex>

#include "rrh_internal.h"

namespace rrh {
namespace {

constexpr int kBlock = 256;

struct GbufOut {
    float4* pos;
    float4* nrm;
    uint32_t* mat;
};

__global__ __launch_bounds__(kBlock) void k_gbuffer(DevScene S, GbufOut o, uint32_t W, uint32_t H) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= W * H) return;
    const rrh_camera& cam = *S.cam;
    const uint32_t px = i % W, py = i / W;
    const f3 d = center_dir(cam, px, py);
    const f3 org = ld3(cam.position);
    const float dd = dot(d, ld3(cam.direction));
    float4 P = make_float4(0.0f, 0.0f, 0.0f, 0.0f), N = P;
    uint32_t M = 0xFFFFFFFFu;
    if (dd > 0.0f) {
        StackPriv st;
        const Hit h = trace_closest4_s<false, true, true>(S, org, d, cam.near_plane / dd, cam.far_plane / dd, st,
                                                           nullptr, nullptr);
        if (h.prim != ~0u) {
            const uint4 tv = S.tri_vtx[h.prim];  // global vertex indices, instance
            const rrh_instance* in = S.insts + tv.w;
            const rrh_mesh* me = S.meshes + in->mesh_index;
            const rrh_vertex* v0 = S.verts + tv.x;
            const rrh_vertex* v1 = S.verts + tv.y;
            const rrh_vertex* v2 = S.verts + tv.z;
            const float b0 = 1.0f - h.b1 - h.b2;
            const f3 w0 = xform_point(in->transform, ld3(v0->pos)), w1 = xform_point(in->transform, ld3(v1->pos)),
                     w2 = xform_point(in->transform, ld3(v2->pos));
            const f3 n0 = normalize(xform_dir(in->transform, ld3(v0->normal))),
                     n1 = normalize(xform_dir(in->transform, ld3(v1->normal))),
                     n2 = normalize(xform_dir(in->transform, ld3(v2->normal)));
            const f3 p = add(add(scale(w0, b0), scale(w1, h.b1)), scale(w2, h.b2));
            const f3 n = normalize(add(add(scale(n0, b0), scale(n1, h.b1)), scale(n2, h.b2)));
            M = me->material_index;
            P = make_float4(p.x, p.y, p.z, 1.0f);
            N = make_float4(n.x, n.y, n.z, __uint_as_float(M));
        }
    }
    o.pos[i] = P;
    o.nrm[i] = N;
    o.mat[i] = M;
}

struct Thresholds {
    float t[256];
};

__device__ __forceinline__ uint32_t tone8(float v, const Thresholds& T) {
    v = (v * (2.51f * v + 0.03f)) / (v * (2.43f * v + 0.59f) + 0.14f);
    v = fminf(fmaxf(v, 0.0f), 1.0f);
    // binary search: the largest k with T_k <= v (T_0 = 0)
    uint32_t k = 0;
#pragma unroll
    for (uint32_t step = 128; step > 0; step >>= 1)
        if (k + step < 256u && T.t[k + step] <= v) k += step;
    return k;
}

template <bool F64>
__global__ __launch_bounds__(kBlock) void k_tonemap(const void* __restrict__ hdr, uint32_t n, float exposure,
                                                    Thresholds T, uint32_t* __restrict__ ldr) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float r, g, b;
    if constexpr (F64) {
        const double* p = static_cast<const double*>(hdr) + (size_t)i * 4;
        r = (float)p[0]; g = (float)p[1]; b = (float)p[2];
    } else {
        const float4 p = static_cast<const float4*>(hdr)[i];
        r = p.x; g = p.y; b = p.z;
    }
    ldr[i] = tone8(r * exposure, T) | tone8(g * exposure, T) << 8 | tone8(b * exposure, T) << 16 | 0xFF000000u;
}

const Thresholds& thresholds() {
    static Thresholds T;
    static std::once_flag once;
    std::call_once(once, [] {
        T.t[0] = 0.0f;
        for (int k = 1; k < 256; ++k) {
            const double t = std::pow(((double)k - 0.5) / 255.0, 2.2);
            float f = (float)t;
            if ((double)f < t) f = std::nextafter(f, INFINITY);
            T.t[k] = f;
        }
    });
    return T;
}

}  // namespace
}  // namespace rrh

using namespace rrh;

extern "C" {

int rrh_gbuffer(rrh_ctx* c, rrh_scene* s, rrh_image* position, rrh_image* normal, rrh_image* material_id) {
    RRH_ARG(c && s && position && normal && material_id, "rrh_gbuffer: null argument");
    RRH_ARG(s->cam_set, "rrh_gbuffer: camera not set (rrh_camera_set)");
    const uint32_t W = s->cam_host.screen_width, H = s->cam_host.screen_height;
    RRH_ARG(position->fmt == RRH_FMT_RGBA32F && normal->fmt == RRH_FMT_RGBA32F && material_id->fmt == RRH_FMT_R32U,
            "rrh_gbuffer: position / normal RGBA32F and material id R32U required");
    RRH_ARG(position->w == W && position->h == H && normal->w == W && normal->h == H && material_id->w == W &&
                material_id->h == H,
            "rrh_gbuffer: image extent differs from the camera");
    RRH_ARG(s->stack4 <= RRH_STACK, "rrh_gbuffer: tree deeper than the traversal stack");
    GbufOut o{(float4*)position->p, (float4*)normal->p, (uint32_t*)material_id->p};
    const uint32_t n = W * H;
    k_gbuffer<<<(n + kBlock - 1) / kBlock, kBlock, 0, c->stream>>>(s->dev(), o, W, H);
    RRH_LAUNCH_CHECK();
    return RRH_OK;
}

int rrh_tonemap(rrh_ctx* c, const rrh_image* hdr, float exposure, rrh_image* ldr) {
    RRH_ARG(c && hdr && ldr, "rrh_tonemap: null argument");
    RRH_ARG(hdr->fmt == RRH_FMT_RGBA32F || hdr->fmt == RRH_FMT_RGBA64F, "rrh_tonemap: RGBA32F or RGBA64F input");
    RRH_ARG(ldr->fmt == RRH_FMT_RGBA8, "rrh_tonemap: RGBA8 output");
    RRH_ARG(hdr->w == ldr->w && hdr->h == ldr->h, "rrh_tonemap: extents differ");
    const uint32_t n = hdr->w * hdr->h;
    const uint32_t g = (n + kBlock - 1) / kBlock;
    if (hdr->fmt == RRH_FMT_RGBA64F)
        k_tonemap<true><<<g, kBlock, 0, c->stream>>>(hdr->p, n, exposure, thresholds(), (uint32_t*)ldr->p);
    else
        k_tonemap<false><<<g, kBlock, 0, c->stream>>>(hdr->p, n, exposure, thresholds(), (uint32_t*)ldr->p);
    RRH_LAUNCH_CHECK();
    return RRH_OK;
}

int rrh_tonemap_thresholds(float out[256]) {
    RRH_ARG(out, "rrh_tonemap_thresholds: null argument");
    const Thresholds& T = thresholds();
    for (int k = 0; k < 256; ++k) out[k] = T.t[k];
    return RRH_OK;
}

}  // extern "C"

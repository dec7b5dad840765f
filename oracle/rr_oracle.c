/*
 * rr_oracle.c -- CPU restatement of the reference renderer's hot path.
 * TEST INFRASTRUCTURE ONLY (see rr_oracle.h).  Built with -ffp-contract=off.
 *
 * Parity status of each function (how it is pinned) is listed in the
 * oracle section of the repository's design notes and in tests/.
 */
#include "rr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ========================================================================== */
/* small vector helpers (source-order fp32)                                   */
/* ========================================================================== */
typedef struct { float x, y, z; } v3;
static inline v3 V(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 vld(const float* p) { return V(p[0], p[1], p[2]); }
static inline v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, v3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 vscale(v3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
static inline v3 vdivs(v3 a, float s) { return V(a.x / s, a.y / s, a.z / s); }
static inline v3 vneg(v3 a) { return V(-a.x, -a.y, -a.z); }
static inline float vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 vcross(v3 a, v3 b) {
    return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float vlength(v3 a) { return sqrtf(vdot(a, a)); }
/* normalize(v) = v * (1 / sqrt(dot(v, v))) */
static inline v3 vnorm(v3 a) { float inv = 1.0f / sqrtf(vdot(a, a)); return vscale(a, inv); }
static inline float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
/* lerp(x, y, s) = x * (1 - s) + y * s */
static inline float lerpf(float x, float y, float s) { return x * (1.0f - s) + y * s; }
static inline float asfloat(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline uint32_t asuint(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline void vst(float* p, v3 a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; }
/* row-major 4x4 times (x, y, z, w), summed left to right */
static inline float mrow(const float* M, int r, float x, float y, float z, float w) {
    return M[r * 4 + 0] * x + M[r * 4 + 1] * y + M[r * 4 + 2] * z + M[r * 4 + 3] * w;
}

/* ========================================================================== */
/* deterministic transcendentals                                              */
/* ========================================================================== */
/* sin/cos: Cody-Waite reduction by pi/2 (three-part split: P1 has 8
 * significant bits, P2 12 bits, so k*P1 and k*P2 are exact for small k),
 * then Taylor series on |r| <= pi/4 (truncation error < 2e-9). */
static void orc_sincosf(float x, float* s, float* c) {
    float k = floorf(x * 0.636619747f + 0.5f);
    float r = x - k * 1.5703125f;
    r = r - k * 4.837512969970703e-4f;
    r = r - k * 7.549790126404332e-8f;
    int q = (int)k;
    float r2 = r * r;
    float sp = r + r * r2 * (-0.166666672f + r2 * (8.33333377e-3f + r2 * (-1.98412701e-4f + r2 * 2.75573188e-6f)));
    float cp = 1.0f + r2 * (-0.5f + r2 * (4.16666679e-2f + r2 * (-1.38888892e-3f + r2 * (2.48015876e-5f + r2 * -2.75573188e-7f))));
    switch (q & 3) {
    case 0: *s = sp; *c = cp; break;
    case 1: *s = cp; *c = -sp; break;
    case 2: *s = -sp; *c = -cp; break;
    default: *s = -cp; *c = sp; break;
    }
}
float orc_sinf(float x) { float s, c; orc_sincosf(x, &s, &c); return s; }
float orc_cosf(float x) { float s, c; orc_sincosf(x, &s, &c); return c; }

/* exp: x = k ln2 + r (two-part ln2), exp(r) by degree-7 Taylor on |r| <= ln2/2,
 * scaled by 2^k built from the exponent bits.  Flushes to 0 below -87. */
float orc_expf(float x) {
    if (x < -87.0f) return 0.0f;
    if (x > 88.0f) x = 88.0f;
    float k = floorf(x * 1.44269502f + 0.5f);
    float r = x - k * 0.693145751953125f;
    r = r - k * 1.42860677e-6f;
    float p = 1.0f + r * (1.0f + r * (0.5f + r * (0.166666672f + r * (4.16666679e-2f + r * (8.33333377e-3f + r * (1.38888892e-3f + r * 1.98412701e-4f))))));
    int ki = (int)k;
    return p * asfloat((uint32_t)(ki + 127) << 23);
}

/* log: x = m 2^e with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s),
 * s = (m-1)/(m+1), odd series to s^9 (|s| < 0.172, error < 1e-9). */
float orc_logf(float x) {
    if (!(x > 0.0f)) return (x == 0.0f) ? -INFINITY : NAN;
    if (isinf(x)) return x;
    int e = 0;
    uint32_t u = asuint(x);
    if (u < 0x00800000u) { u = asuint(x * 8388608.0f); e = -23; }
    e += (int)((u >> 23) & 255u) - 127;
    float m = asfloat((u & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421354f) { m = m * 0.5f; e = e + 1; }
    float s = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float p = s * (2.0f + s2 * (0.666666687f + s2 * (0.400000006f + s2 * (0.285714298f + s2 * 0.222222224f))));
    float ef = (float)e;
    return ef * 0.693145751953125f + (p + ef * 1.42860677e-6f);
}

/* ========================================================================== */
/* RNG -- Common.slang:18-28, Sampling.slang:11-30                            */
/* ========================================================================== */
uint32_t orc_pcg_hash(uint32_t seed) {
    uint32_t state = seed * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
float orc_uint_to_float01(uint32_t v) { return (float)(v >> 8) * (1.0f / 16777216.0f); }
float orc_rand_float(uint32_t* seed) { *seed = orc_pcg_hash(*seed); return orc_uint_to_float01(*seed); }
uint32_t orc_make_seed(uint32_t px, uint32_t py, uint32_t frame) {
    return orc_pcg_hash(px ^ (py * 1000003u) ^ (frame * 2654435761u));
}

/* ========================================================================== */
/* camera -- Camera.cpp:27-36,120-146 (glm lookAt / perspective_NO / inverse),
 * GpuCameraData fill Scene.cpp:554-585.  Evaluated in double, stored fp32.  */
/* ========================================================================== */
static void inv4d(const double* m, double* out) {
    double inv[16], det;
    inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
    det = 1.0 / det;
    for (int i = 0; i < 16; ++i) out[i] = inv[i] * det;
}
static void mulmat4d(const double* a, const double* b, double* o) {
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            double s = 0;
            for (int k = 0; k < 4; ++k) s += a[r * 4 + k] * b[k * 4 + c];
            o[r * 4 + c] = s;
        }
}

void orc_camera_look_at(const float eye[3], const float center[3], float fov_y_deg,
                        uint32_t width, uint32_t height, float near_plane, float far_plane,
                        uint32_t frame_index, orc_camera* out) {
    const double PI = 3.14159265358979323846;
    double d[3] = {(double)center[0] - eye[0], (double)center[1] - eye[1], (double)center[2] - eye[2]};
    double dl = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int i = 0; i < 3; ++i) d[i] /= dl;
    /* Camera::look_at keeps yaw/pitch; forward() rebuilds the direction */
    double pitch = asin(d[1]), yaw = atan2(d[2], d[0]);
    double f[3] = {cos(pitch) * cos(yaw), sin(pitch), cos(pitch) * sin(yaw)};
    double fl = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (int i = 0; i < 3; ++i) f[i] /= fl;
    /* right = normalize(cross(fwd, (0,1,0))); up = normalize(cross(right, fwd)) */
    double r[3] = {-f[2], 0.0, f[0]};
    double rl = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    if (rl < 1e-5) { r[0] = 1; r[1] = 0; r[2] = 0; rl = 1; }
    for (int i = 0; i < 3; ++i) r[i] /= rl;
    double u[3] = {r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0]};
    double ul = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    for (int i = 0; i < 3; ++i) u[i] /= ul;
    /* glm::lookAtRH(eye, eye + f, up): s = normalize(cross(f, up)), u' = cross(s, f) */
    double s[3] = {f[1] * u[2] - f[2] * u[1], f[2] * u[0] - f[0] * u[2], f[0] * u[1] - f[1] * u[0]};
    double sl = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (int i = 0; i < 3; ++i) s[i] /= sl;
    double uu[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
    double e[3] = {eye[0], eye[1], eye[2]};
    double view[16] = {
        s[0], s[1], s[2], -(s[0] * e[0] + s[1] * e[1] + s[2] * e[2]),
        uu[0], uu[1], uu[2], -(uu[0] * e[0] + uu[1] * e[1] + uu[2] * e[2]),
        -f[0], -f[1], -f[2], (f[0] * e[0] + f[1] * e[1] + f[2] * e[2]),
        0, 0, 0, 1};
    double aspect = (double)width / (double)height;
    double th = tan((double)fov_y_deg * PI / 180.0 * 0.5);
    double n = near_plane, fr = far_plane;
    double proj[16] = {0};
    proj[0] = 1.0 / (aspect * th);
    proj[5] = -1.0 / th; /* Vulkan Y flip, Camera.cpp:142 */
    proj[10] = -(fr + n) / (fr - n);
    proj[11] = -(2.0 * fr * n) / (fr - n);
    proj[14] = -1.0;
    double iv[16], ip[16], vp[16];
    inv4d(view, iv);
    inv4d(proj, ip);
    mulmat4d(proj, view, vp);
    memset(out, 0, sizeof(*out));
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
}

/* pathtracer.slang:195-205 (also restir_di.slang:199-206) */
void orc_primary_ray(const orc_camera* cam, uint32_t px, uint32_t py, uint32_t* seed,
                     float org[3], float dir[3]) {
    float u0 = orc_rand_float(seed);
    float u1 = orc_rand_float(seed);
    float ndcx = ((float)px + u0) / (float)cam->screen_width * 2.0f - 1.0f;
    float ndcy = ((float)py + u1) / (float)cam->screen_height * 2.0f - 1.0f;
    float vx = mrow(cam->inv_proj, 0, ndcx, ndcy, 1.0f, 1.0f);
    float vy = mrow(cam->inv_proj, 1, ndcx, ndcy, 1.0f, 1.0f);
    v3 w = V(mrow(cam->inv_view, 0, vx, vy, -1.0f, 0.0f),
             mrow(cam->inv_view, 1, vx, vy, -1.0f, 0.0f),
             mrow(cam->inv_view, 2, vx, vy, -1.0f, 0.0f));
    vst(dir, vnorm(w));
    org[0] = cam->position[0]; org[1] = cam->position[1]; org[2] = cam->position[2];
}

/* pixel-centre sky direction, restir_di.slang:355-358 */
static v3 center_ray_dir(const orc_camera* cam, uint32_t px, uint32_t py) {
    float ndcx = ((float)px + 0.5f) / (float)cam->screen_width * 2.0f - 1.0f;
    float ndcy = ((float)py + 0.5f) / (float)cam->screen_height * 2.0f - 1.0f;
    float vx = mrow(cam->inv_proj, 0, ndcx, ndcy, 1.0f, 1.0f);
    float vy = mrow(cam->inv_proj, 1, ndcx, ndcy, 1.0f, 1.0f);
    v3 w = V(mrow(cam->inv_view, 0, vx, vy, -1.0f, 0.0f),
             mrow(cam->inv_view, 1, vx, vy, -1.0f, 0.0f),
             mrow(cam->inv_view, 2, vx, vy, -1.0f, 0.0f));
    return vnorm(w);
}

/* sky gradient, pathtracer.slang:215-217 */
static v3 sky(v3 d) {
    float t = 0.5f * (d.y + 1.0f);
    return V(lerpf(0.1f, 0.2f, t), lerpf(0.1f, 0.3f, t), lerpf(0.1f, 0.5f, t));
}

/* ========================================================================== */
/* scene: flattening (Scene.cpp:255-515 upload + per-instance transform as in
 * pathtracer.slang:103) and the oracle's own BVH                            */
/* ========================================================================== */
typedef struct { float lo[3], hi[3]; int32_t left, right; uint32_t first, count; } obvh_node;

struct orc_scene {
    const orc_vertex* v; uint32_t nv;
    const uint32_t* idx; uint32_t ni;
    const orc_mesh* m; uint32_t nm;
    const orc_material* mat; uint32_t nmat;
    const orc_instance* inst; uint32_t ninst;
    const orc_light* l; uint32_t nl;
    uint32_t flags;
    uint32_t ntri;
    float* tri;           /* ntri x 9 world-space vertices */
    uint32_t* tri_inst;   /* instance of each flattened triangle */
    uint32_t* tri_prim;   /* primitive index within the mesh */
    uint32_t* tri_light_of;  /* per tri: light index or ~0u */
    uint32_t ntl;         /* triangle lights */
    uint32_t* tl_tri;     /* triangle id of each triangle light */
    obvh_node* nodes; uint32_t nnodes;
    uint32_t* order;      /* tri ids in leaf order */
};

static v3 xform_point(const float* T, v3 p) {
    return V(mrow(T, 0, p.x, p.y, p.z, 1.0f), mrow(T, 1, p.x, p.y, p.z, 1.0f), mrow(T, 2, p.x, p.y, p.z, 1.0f));
}
static v3 xform_dir(const float* T, v3 p) {
    return V(T[0] * p.x + T[1] * p.y + T[2] * p.z, T[4] * p.x + T[5] * p.y + T[6] * p.z,
             T[8] * p.x + T[9] * p.y + T[10] * p.z);
}

static int emissive(const orc_material* m) {
    return m->emissive[0] > 0.0f || m->emissive[1] > 0.0f || m->emissive[2] > 0.0f;
}

static uint32_t bvh_build_rec(orc_scene* s, const float* cen, uint32_t first, uint32_t count, uint32_t* cap) {
    uint32_t ni = s->nnodes++;
    if (s->nnodes > *cap) { *cap *= 2; s->nodes = (obvh_node*)realloc(s->nodes, *cap * sizeof(obvh_node)); }
    obvh_node nd;
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < 3; ++k) { nd.lo[k] = INFINITY; nd.hi[k] = -INFINITY; }
    for (uint32_t i = first; i < first + count; ++i) {
        uint32_t t = s->order[i];
        for (int c = 0; c < 3; ++c) {
            for (int k = 0; k < 3; ++k) {
                float x = s->tri[t * 9 + c * 3 + k];
                nd.lo[k] = fminf(nd.lo[k], x); nd.hi[k] = fmaxf(nd.hi[k], x);
            }
        }
        for (int k = 0; k < 3; ++k) { clo[k] = fminf(clo[k], cen[t * 3 + k]); chi[k] = fmaxf(chi[k], cen[t * 3 + k]); }
    }
    nd.left = nd.right = -1; nd.first = first; nd.count = count;
    if (count > 4) {
        int ax = 0;
        float ext[3] = {chi[0] - clo[0], chi[1] - clo[1], chi[2] - clo[2]};
        if (ext[1] > ext[ax]) ax = 1;
        if (ext[2] > ext[ax]) ax = 2;
        if (ext[ax] > 0.0f) {
            /* partition by centroid midpoint; fall back to an index split */
            float mid = 0.5f * (clo[ax] + chi[ax]);
            uint32_t i = first, j = first + count;
            while (i < j) {
                if (cen[s->order[i] * 3 + ax] < mid) ++i;
                else { --j; uint32_t t = s->order[i]; s->order[i] = s->order[j]; s->order[j] = t; }
            }
            uint32_t nl = i - first;
            if (nl == 0 || nl == count) nl = count / 2;
            s->nodes[ni] = nd;
            uint32_t L = bvh_build_rec(s, cen, first, nl, cap);
            uint32_t R = bvh_build_rec(s, cen, first + nl, count - nl, cap);
            s->nodes[ni].left = (int32_t)L; s->nodes[ni].right = (int32_t)R; s->nodes[ni].count = 0;
            return ni;
        }
    }
    s->nodes[ni] = nd;
    return ni;
}

orc_scene* orc_scene_create(const orc_vertex* v, uint32_t nv, const uint32_t* idx, uint32_t ni,
                            const orc_mesh* m, uint32_t nm, const orc_material* mat, uint32_t nmat,
                            const orc_instance* inst, uint32_t ninst, const orc_light* l, uint32_t nl,
                            uint32_t flags) {
    orc_scene* s = (orc_scene*)calloc(1, sizeof(orc_scene));
    s->v = v; s->nv = nv; s->idx = idx; s->ni = ni; s->m = m; s->nm = nm; s->mat = mat; s->nmat = nmat;
    s->inst = inst; s->ninst = ninst; s->l = l; s->nl = nl; s->flags = flags;
    uint32_t ntri = 0;
    for (uint32_t i = 0; i < ninst; ++i) ntri += m[inst[i].mesh_index].index_count / 3;
    s->ntri = ntri;
    s->tri = (float*)malloc((size_t)ntri * 9 * sizeof(float));
    s->tri_inst = (uint32_t*)malloc((size_t)ntri * 4);
    s->tri_prim = (uint32_t*)malloc((size_t)ntri * 4);
    s->tri_light_of = (uint32_t*)malloc((size_t)ntri * 4);
    s->tl_tri = (uint32_t*)malloc((size_t)(ntri ? ntri : 1) * 4);
    uint32_t t = 0;
    for (uint32_t i = 0; i < ninst; ++i) {
        const orc_mesh* me = &m[inst[i].mesh_index];
        int em = emissive(&mat[me->material_index]);
        for (uint32_t p = 0; p < me->index_count / 3; ++p, ++t) {
            uint32_t base = me->index_byte_offset / 4 + p * 3;
            for (int c = 0; c < 3; ++c) {
                const orc_vertex* vv = &v[me->vertex_base + idx[base + c]];
                vst(&s->tri[(size_t)t * 9 + c * 3], xform_point(inst[i].transform, vld(vv->pos)));
            }
            s->tri_inst[t] = i;
            s->tri_prim[t] = p;
            s->tri_light_of[t] = ~0u;
            if ((flags & ORC_SCENE_TRI_LIGHTS) && em) {
                s->tri_light_of[t] = nl + s->ntl;
                s->tl_tri[s->ntl++] = t;
            }
        }
    }
    /* oracle BVH */
    float* cen = (float*)malloc((size_t)(ntri ? ntri : 1) * 3 * sizeof(float));
    s->order = (uint32_t*)malloc((size_t)(ntri ? ntri : 1) * 4);
    for (uint32_t i = 0; i < ntri; ++i) {
        s->order[i] = i;
        for (int k = 0; k < 3; ++k)
            cen[i * 3 + k] = (s->tri[i * 9 + k] + s->tri[i * 9 + 3 + k] + s->tri[i * 9 + 6 + k]) * (1.0f / 3.0f);
    }
    uint32_t cap = 1024;
    s->nodes = (obvh_node*)malloc(cap * sizeof(obvh_node));
    if (ntri) bvh_build_rec(s, cen, 0, ntri, &cap);
    free(cen);
    return s;
}

void orc_scene_destroy(orc_scene* s) {
    if (!s) return;
    free(s->tri); free(s->tri_inst); free(s->tri_prim); free(s->tri_light_of); free(s->tl_tri);
    free(s->nodes); free(s->order); free(s);
}
uint32_t orc_scene_tri_count(const orc_scene* s) { return s->ntri; }
uint32_t orc_scene_light_count(const orc_scene* s) { return s->nl + s->ntl; }
void orc_scene_triangles(const orc_scene* s, float* out) { memcpy(out, s->tri, (size_t)s->ntri * 36); }

/* ========================================================================== */
/* watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2(1) 2013), two-
 * sided (the reference instances are CULL_DISABLE, AccelStructure.cpp:90)   */
/* ========================================================================== */
typedef struct { int kx, ky, kz; float Sx, Sy, Sz; v3 o; v3 d; v3 inv; } wray;

static void wray_setup(v3 o, v3 d, wray* r) {
    float a[3] = {fabsf(d.x), fabsf(d.y), fabsf(d.z)};
    int kz = (a[0] > a[1]) ? ((a[0] > a[2]) ? 0 : 2) : ((a[1] > a[2]) ? 1 : 2);
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    float dd[3] = {d.x, d.y, d.z};
    if (dd[kz] < 0.0f) { int t = kx; kx = ky; ky = t; }
    r->kx = kx; r->ky = ky; r->kz = kz;
    r->Sx = dd[kx] / dd[kz];
    r->Sy = dd[ky] / dd[kz];
    r->Sz = 1.0f / dd[kz];
    r->o = o; r->d = d;
    r->inv = V(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
}

static inline float comp(v3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

/* returns 1 and (t, b1, b2) when the hit lies in [tmin, tmax] */
static int wt_hit(const wray* r, const float* tv, float tmin, float tmax, float* t, float* b1, float* b2) {
    v3 A = vsub(vld(tv), r->o), B = vsub(vld(tv + 3), r->o), C = vsub(vld(tv + 6), r->o);
    float Az = comp(A, r->kz), Bz = comp(B, r->kz), Cz = comp(C, r->kz);
    float Ax = comp(A, r->kx) - r->Sx * Az, Ay = comp(A, r->ky) - r->Sy * Az;
    float Bx = comp(B, r->kx) - r->Sx * Bz, By = comp(B, r->ky) - r->Sy * Bz;
    float Cx = comp(C, r->kx) - r->Sx * Cz, Cy = comp(C, r->ky) - r->Sy * Cz;
    float U = Cx * By - Cy * Bx;
    float Vv = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || Vv == 0.0f || W == 0.0f) {
        /* edge case: recompute the edge functions in double (paper, section 4) */
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        Vv = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || Vv < 0.0f || W < 0.0f) && (U > 0.0f || Vv > 0.0f || W > 0.0f)) return 0;
    float det = U + Vv + W;
    if (det == 0.0f) return 0;
    Az = r->Sz * Az; Bz = r->Sz * Bz; Cz = r->Sz * Cz;
    float T = U * Az + Vv * Bz + W * Cz;
    float rcp = 1.0f / det;
    float tt = T * rcp;
    if (!(tt >= tmin && tt <= tmax)) return 0;
    *t = tt; *b1 = Vv * rcp; *b2 = W * rcp;
    return 1;
}

/* conservative slab test: far distance scaled by 1 + 2*gamma(3) (PBRT-v3 3.9.2) */
static int box_hit(const float* lo, const float* hi, const wray* r, float tmin, float tmax) {
    float t0x = (lo[0] - r->o.x) * r->inv.x, t1x = (hi[0] - r->o.x) * r->inv.x;
    float t0y = (lo[1] - r->o.y) * r->inv.y, t1y = (hi[1] - r->o.y) * r->inv.y;
    float t0z = (lo[2] - r->o.z) * r->inv.z, t1z = (hi[2] - r->o.z) * r->inv.z;
    float tn = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), tmin));
    /* tmax is the current closest t: widened as well so an exact tie (equal t, smaller id) is never culled */
    float tf = fminf(fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z)) * 1.0000008f, tmax * 1.000004f);
    return tn <= tf;
}

/* closest-hit rule shared with the GPU kernels: smaller t wins, equal t -> smaller id */
static inline int better(float t, uint32_t prim, float bt, uint32_t bp) {
    return t < bt || (t == bt && prim < bp);
}

/* a triangle faces the ray when its geometric normal (v1 - v0) x (v2 - v0) points against it:
 * the counter-clockwise front face of the reference's G-buffer raster pipeline
 * (GBufferPass.cpp:148-151, Pipeline.h:75) */
static int front_facing(const float* tv, v3 d) {
    v3 p0 = vld(tv), p1 = vld(tv + 3), p2 = vld(tv + 6);
    return vdot(vcross(vsub(p1, p0), vsub(p2, p0)), d) < 0.0f;
}

static orc_hit trace_bvh_cull(const orc_scene* s, v3 o, v3 d, float tmin, float tmax, int cull);
static orc_hit trace_bvh_one(const orc_scene* s, v3 o, v3 d, float tmin, float tmax) {
    return trace_bvh_cull(s, o, d, tmin, tmax, 0);
}
static orc_hit trace_bvh_cull(const orc_scene* s, v3 o, v3 d, float tmin, float tmax, int cull) {
    orc_hit h; h.t = tmax; h.prim = ~0u; h.b1 = h.b2 = 0.0f;
    if (s->ntri == 0) return h;
    wray r; wray_setup(o, d, &r);
    uint32_t stack[128]; int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const obvh_node* nd = &s->nodes[stack[--sp]];
        if (!box_hit(nd->lo, nd->hi, &r, tmin, h.t)) continue;
        if (nd->left < 0) {
            for (uint32_t i = nd->first; i < nd->first + nd->count; ++i) {
                uint32_t t = s->order[i];
                float tt, b1, b2;
                if (wt_hit(&r, &s->tri[(size_t)t * 9], tmin, h.t, &tt, &b1, &b2) && better(tt, t, h.t, h.prim) &&
                    (!cull || front_facing(&s->tri[(size_t)t * 9], d))) {
                    h.t = tt; h.prim = t; h.b1 = b1; h.b2 = b2;
                }
            }
        } else {
            stack[sp++] = (uint32_t)nd->right;
            stack[sp++] = (uint32_t)nd->left;
        }
    }
    return h;
}

static int occluded_bvh(const orc_scene* s, v3 o, v3 d, float tmin, float tmax) {
    if (s->ntri == 0 || !(tmax >= tmin)) return 0;
    wray r; wray_setup(o, d, &r);
    uint32_t stack[128]; int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const obvh_node* nd = &s->nodes[stack[--sp]];
        if (!box_hit(nd->lo, nd->hi, &r, tmin, tmax)) continue;
        if (nd->left < 0) {
            for (uint32_t i = nd->first; i < nd->first + nd->count; ++i) {
                float tt, b1, b2;
                if (wt_hit(&r, &s->tri[(size_t)s->order[i] * 9], tmin, tmax, &tt, &b1, &b2)) return 1;
            }
        } else {
            stack[sp++] = (uint32_t)nd->right;
            stack[sp++] = (uint32_t)nd->left;
        }
    }
    return 0;
}

int orc_occluded(const orc_scene* s, const float o[3], const float d[3], float tmin, float tmax) {
    return occluded_bvh(s, vld(o), vld(d), tmin, tmax);
}

void orc_trace_brute(const orc_scene* s, const float* rays, uint32_t n, orc_hit* hits) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const float* ry = rays + i * 8;
        wray r; wray_setup(vld(ry), vld(ry + 4), &r);
        orc_hit h; h.t = ry[7]; h.prim = ~0u; h.b1 = h.b2 = 0.0f;
        for (uint32_t t = 0; t < s->ntri; ++t) {
            float tt, b1, b2;
            if (wt_hit(&r, &s->tri[(size_t)t * 9], ry[3], h.t, &tt, &b1, &b2) && better(tt, t, h.t, h.prim)) {
                h.t = tt; h.prim = t; h.b1 = b1; h.b2 = b2;
            }
        }
        hits[i] = h;
    }
}

void orc_trace_bvh(const orc_scene* s, const float* rays, uint32_t n, orc_hit* hits) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const float* ry = rays + i * 8;
        hits[i] = trace_bvh_one(s, vld(ry), vld(ry + 4), ry[3], ry[7]);
    }
}

/* hit attributes exactly as pathtracer.slang:69-117 fetches them */
typedef struct { v3 pos, normal; uint32_t mat; float t; uint32_t prim; int hit; } hitinfo;

static hitinfo fetch_hit(const orc_scene* s, orc_hit h) {
    hitinfo hi; memset(&hi, 0, sizeof(hi));
    hi.hit = (h.prim != ~0u);
    hi.t = h.t; hi.prim = h.prim;
    if (!hi.hit) return hi;
    const orc_instance* in = &s->inst[s->tri_inst[h.prim]];
    const orc_mesh* me = &s->m[in->mesh_index];
    uint32_t base = me->index_byte_offset / 4 + s->tri_prim[h.prim] * 3;
    const orc_vertex* v0 = &s->v[me->vertex_base + s->idx[base]];
    const orc_vertex* v1 = &s->v[me->vertex_base + s->idx[base + 1]];
    const orc_vertex* v2 = &s->v[me->vertex_base + s->idx[base + 2]];
    float b0 = 1.0f - h.b1 - h.b2;
    v3 lp = vadd(vadd(vscale(vld(v0->pos), b0), vscale(vld(v1->pos), h.b1)), vscale(vld(v2->pos), h.b2));
    v3 ln = vnorm(vadd(vadd(vscale(vld(v0->normal), b0), vscale(vld(v1->normal), h.b1)), vscale(vld(v2->normal), h.b2)));
    hi.pos = xform_point(in->transform, lp);
    hi.normal = vnorm(xform_dir(in->transform, ln));
    hi.mat = me->material_index;
    return hi;
}

static hitinfo trace(const orc_scene* s, v3 o, v3 d, float tmin, float tmax) {
    return fetch_hit(s, trace_bvh_one(s, o, d, tmin, tmax));
}

/* G-buffer of the reference's raster pass (gbuffer.slang vs_main / fs_main, GBufferPass.cpp:
 * 216-330) by one ray through each pixel centre: front faces only, depth between the near and
 * far planes (view depth = t * dot(dir, camera forward)); position = barycentric blend of
 * the transformed vertices, w = 1; normal = normalize(blend of normalize(M3 n_i)), w =
 * asfloat(material); material id; a miss keeps the clear values (0, 0, 0xFFFFFFFF). */
void orc_gbuffer(const orc_scene* s, const orc_camera* cam, float* pos, float* nrm, uint32_t* mat) {
    const uint32_t W = cam->screen_width, H = cam->screen_height;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t py = 0; py < (int64_t)H; ++py)
        for (uint32_t px = 0; px < W; ++px) {
            const size_t i = (size_t)py * W + px;
            v3 d = center_ray_dir(cam, px, (uint32_t)py);
            v3 o = vld(cam->position);
            float dd = vdot(d, vld(cam->direction));
            float* P = pos + i * 4;
            float* N = nrm + i * 4;
            P[0] = P[1] = P[2] = P[3] = 0.0f;
            N[0] = N[1] = N[2] = N[3] = 0.0f;
            mat[i] = 0xFFFFFFFFu;
            if (!(dd > 0.0f)) continue;
            orc_hit h = trace_bvh_cull(s, o, d, cam->near_plane / dd, cam->far_plane / dd, 1);
            if (h.prim == ~0u) continue;
            const orc_instance* in = &s->inst[s->tri_inst[h.prim]];
            const orc_mesh* me = &s->m[in->mesh_index];
            uint32_t base = me->index_byte_offset / 4 + s->tri_prim[h.prim] * 3;
            const orc_vertex* v0 = &s->v[me->vertex_base + s->idx[base]];
            const orc_vertex* v1 = &s->v[me->vertex_base + s->idx[base + 1]];
            const orc_vertex* v2 = &s->v[me->vertex_base + s->idx[base + 2]];
            float b0 = 1.0f - h.b1 - h.b2;
            v3 w0 = xform_point(in->transform, vld(v0->pos)), w1 = xform_point(in->transform, vld(v1->pos)),
               w2 = xform_point(in->transform, vld(v2->pos));
            v3 n0 = vnorm(xform_dir(in->transform, vld(v0->normal))), n1 = vnorm(xform_dir(in->transform, vld(v1->normal))),
               n2 = vnorm(xform_dir(in->transform, vld(v2->normal)));
            v3 p = vadd(vadd(vscale(w0, b0), vscale(w1, h.b1)), vscale(w2, h.b2));
            v3 n = vnorm(vadd(vadd(vscale(n0, b0), vscale(n1, h.b1)), vscale(n2, h.b2)));
            vst(P, p); P[3] = 1.0f;
            vst(N, n); N[3] = asfloat(me->material_index);
            mat[i] = me->material_index;
        }
}

/* the screenshot's 8-bit values, Application.cpp:853-866 as written: ACES (Narkowicz) on the
 * exposed value, clamp to [0, 1], powf(v, 1 / 2.2), (uint8)(v * 255 + 0.5); a NaN counts as 0 */
void orc_tonemap(const float* hdr, size_t pixels, float exposure, uint8_t* ldr) {
    for (size_t i = 0; i < pixels; ++i) {
        for (int c = 0; c < 3; ++c) {
            float v = hdr[i * 4 + c] * exposure;
            v = (v * (2.51f * v + 0.03f)) / (v * (2.43f * v + 0.59f) + 0.14f);
            v = fminf(fmaxf(v, 0.0f), 1.0f);
            v = powf(v, 1.0f / 2.2f);
            ldr[i * 4 + c] = (uint8_t)(v * 255.0f + 0.5f);
        }
        ldr[i * 4 + 3] = 255u;
    }
}

/* ========================================================================== */
/* sampling + BRDF -- Sampling.slang:37-66, BRDF.slang:16-187                  */
/* ========================================================================== */
#define K_PI 3.14159265358979323846f
#define K_TWO_PI 6.28318530717958647692f
#define K_INV_PI 0.31830988618379067154f
#define K_EPS 1e-4f

void orc_sample_cosine_hemisphere(float u0, float u1, float out[3], float* pdf) {
    float r = sqrtf(u0);
    float phi = K_TWO_PI * u1;
    float s, c; orc_sincosf(phi, &s, &c);
    float x = r * c, y = r * s;
    float z = sqrtf(fmaxf(0.0f, 1.0f - u0));
    *pdf = z * K_INV_PI;
    out[0] = x; out[1] = y; out[2] = z;
}

static void build_onb(v3 n, v3* T, v3* B) {
    v3 up = (fabsf(n.x) < 0.9f) ? V(1.0f, 0.0f, 0.0f) : V(0.0f, 1.0f, 0.0f);
    *T = vnorm(vcross(up, n));
    *B = vcross(n, *T);
}
void orc_build_onb(const float n[3], float T[3], float B[3]) {
    v3 t, b; build_onb(vld(n), &t, &b); vst(T, t); vst(B, b);
}
static v3 to_world(v3 l, v3 T, v3 B, v3 N) {
    return vadd(vadd(vscale(T, l.x), vscale(B, l.y)), vscale(N, l.z));
}

static v3 fresnel_schlick(float cos_theta, v3 F0) {
    float x = clampf(1.0f - cos_theta, 0.0f, 1.0f);
    float x2 = x * x;
    float p = x2 * x2 * x;
    return V(F0.x + (1.0f - F0.x) * p, F0.y + (1.0f - F0.y) * p, F0.z + (1.0f - F0.z) * p);
}
static float distribution_ggx(float NdotH, float alpha) {
    float a2 = alpha * alpha;
    float d = NdotH * NdotH * (a2 - 1.0f) + 1.0f;
    return a2 / fmaxf(K_PI * d * d, K_EPS);
}
static float geometry_smith_ggx(float NdotV, float NdotL, float alpha) {
    float k = alpha * 0.5f;
    float gv = NdotV / fmaxf(NdotV * (1.0f - k) + k, K_EPS);
    float gl = NdotL / fmaxf(NdotL * (1.0f - k) + k, K_EPS);
    return gv * gl;
}
static v3 brdf_specular_ggx(v3 n, v3 wo, v3 wi, v3 F0, float roughness) {
    float alpha = roughness * roughness;
    v3 h = vnorm(vadd(wo, wi));
    float NdotH = fmaxf(vdot(n, h), 0.0f);
    float NdotV = fmaxf(vdot(n, wo), 0.0f);
    float NdotL = fmaxf(vdot(n, wi), 0.0f);
    float VdotH = fmaxf(vdot(wo, h), 0.0f);
    v3 F = fresnel_schlick(VdotH, F0);
    float D = distribution_ggx(NdotH, alpha);
    float G = geometry_smith_ggx(NdotV, NdotL, alpha);
    float den = fmaxf(4.0f * NdotV * NdotL, K_EPS);
    float DG = D * G;
    return V(DG * F.x / den, DG * F.y / den, DG * F.z / den);
}
static v3 F0_of(v3 base, float metallic) {
    return V(lerpf(0.04f, base.x, metallic), lerpf(0.04f, base.y, metallic), lerpf(0.04f, base.z, metallic));
}
static v3 brdf_eval(v3 n, v3 wo, v3 wi, v3 base, float metallic, float roughness) {
    float r = clampf(roughness, 0.04f, 1.0f);
    v3 F0 = F0_of(base, metallic);
    v3 kS = fresnel_schlick(fmaxf(vdot(n, wo), 0.0f), F0);
    v3 kD = V((1.0f - kS.x) * (1.0f - metallic), (1.0f - kS.y) * (1.0f - metallic), (1.0f - kS.z) * (1.0f - metallic));
    v3 diffuse = vmul(kD, vscale(base, K_INV_PI));
    v3 spec = brdf_specular_ggx(n, wo, wi, F0, r);
    return vadd(diffuse, spec);
}
void orc_brdf_eval(const float n[3], const float wo[3], const float wi[3],
                   const float base[3], float metallic, float roughness, float out[3]) {
    vst(out, brdf_eval(vld(n), vld(wo), vld(wi), vld(base), metallic, roughness));
}

static v3 sample_ggx_vndf(v3 wo, float roughness, float u0, float u1) {
    float alpha = roughness * roughness;
    v3 Vh = vnorm(V(alpha * wo.x, alpha * wo.y, wo.z));
    float lensq = Vh.x * Vh.x + Vh.y * Vh.y;
    v3 T1;
    if (lensq > 0.0f) { float il = sqrtf(lensq); T1 = V(-Vh.y / il, Vh.x / il, 0.0f / il); }
    else T1 = V(1.0f, 0.0f, 0.0f);
    v3 T2 = vcross(Vh, T1);
    float r2 = sqrtf(u0);
    float phi = K_TWO_PI * u1;
    float s, c; orc_sincosf(phi, &s, &c);
    float t1 = r2 * c, t2 = r2 * s;
    float sv = 0.5f * (1.0f + Vh.z);
    t2 = (1.0f - sv) * sqrtf(1.0f - t1 * t1) + sv * t2;
    float tz = sqrtf(fmaxf(0.0f, 1.0f - t1 * t1 - t2 * t2));
    v3 Nh = vadd(vadd(vscale(T1, t1), vscale(T2, t2)), vscale(Vh, tz));
    return vnorm(V(alpha * Nh.x, alpha * Nh.y, fmaxf(0.0f, Nh.z)));
}

static float p_spec_of(v3 n, v3 wo, v3 base, float metallic) {
    v3 kS = fresnel_schlick(fmaxf(vdot(n, wo), 0.0f), F0_of(base, metallic));
    float p = (kS.x + kS.y + kS.z) / 3.0f;
    return clampf(p, 0.1f, 0.9f);
}

/* BRDF.slang:116-172 (returns f and the mixture pdf; f = 0, pdf = 1 below the horizon) */
static v3 brdf_sample(v3 wo, v3 n, v3 base, float metallic, float roughness, float ux, float uy,
                      v3* wi, float* pdf) {
    float r = clampf(roughness, 0.04f, 1.0f);
    float p_spec = p_spec_of(n, wo, base, metallic);
    v3 T, B; build_onb(n, &T, &B);
    v3 wol = V(vdot(wo, T), vdot(wo, B), vdot(wo, n));
    v3 wil; float pdf_local;
    if (ux < p_spec) {
        v3 h = sample_ggx_vndf(wol, r, ux / p_spec, uy);
        float d2 = 2.0f * vdot(h, vneg(wol));
        wil = vsub(vneg(wol), vscale(h, d2)); /* reflect(-wo, h) = i - 2 dot(h, i) h */
        if (wil.z <= 0.0f) { *wi = V(0.0f, 1.0f, 0.0f); *pdf = 1.0f; return V(0.0f, 0.0f, 0.0f); }
    } else {
        float c[3];
        orc_sample_cosine_hemisphere((ux - p_spec) / (1.0f - p_spec), uy, c, &pdf_local);
        wil = vld(c);
    }
    *wi = to_world(wil, T, B, n);
    float NdotWI = fmaxf(wil.z, 0.0f);
    v3 h = vnorm(vadd(wol, wil));
    float NdotH = fmaxf(h.z, 1e-6f);
    float VdotH = fmaxf(vdot(wol, h), 1e-6f);
    float D = distribution_ggx(NdotH, r * r);
    float pdf_spec = D * NdotH / fmaxf(4.0f * VdotH, 1e-6f);
    float pdf_diff = NdotWI * K_INV_PI;
    float p = p_spec * pdf_spec + (1.0f - p_spec) * pdf_diff;
    *pdf = fmaxf(p, 1e-6f);
    return brdf_eval(n, wo, *wi, base, metallic, roughness);
}
void orc_brdf_sample(const float wo[3], const float n[3], const float base[3], float metallic,
                     float roughness, float u0, float u1, float wi[3], float* pdf, float f[3]) {
    v3 w; vst(f, brdf_sample(vld(wo), vld(n), vld(base), metallic, roughness, u0, u1, &w, pdf));
    vst(wi, w);
}

/* BRDF.slang:176-187 (fixed 0.5/0.5 mixture) */
static float brdf_pdf(v3 wo, v3 wi, v3 n, float roughness) {
    float r = clampf(roughness, 0.04f, 1.0f);
    v3 h = vnorm(vadd(wo, wi));
    float NdotH = fmaxf(vdot(n, h), 1e-6f);
    float VdotH = fmaxf(vdot(wo, h), 1e-6f);
    float NdotL = fmaxf(vdot(n, wi), 0.0f);
    float D = distribution_ggx(NdotH, r * r);
    float pdf_spec = D * NdotH / fmaxf(4.0f * VdotH, 1e-6f);
    float pdf_diff = NdotL * K_INV_PI;
    return 0.5f * pdf_spec + 0.5f * pdf_diff;
}
float orc_brdf_pdf(const float wo[3], const float wi[3], const float n[3], float roughness) {
    return brdf_pdf(vld(wo), vld(wi), vld(n), roughness);
}

/* the BRDF.slang pieces one at a time and over arrays, so tests can integrate and histogram
 * them against closed forms (no reference counterpart: checking entry points only) */
void orc_fresnel_schlick(float cos_theta, const float F0[3], float out[3]) {
    vst(out, fresnel_schlick(cos_theta, vld(F0)));
}
float orc_distribution_ggx(float NdotH, float alpha) { return distribution_ggx(NdotH, alpha); }
float orc_geometry_smith_ggx(float NdotV, float NdotL, float alpha) {
    return geometry_smith_ggx(NdotV, NdotL, alpha);
}
void orc_sample_ggx_vndf_n(const float wo_local[3], float roughness, const float* u, uint32_t count, float* h) {
    v3 wol = vld(wo_local);
    for (uint32_t i = 0; i < count; ++i) vst(h + 3 * i, sample_ggx_vndf(wol, roughness, u[2 * i], u[2 * i + 1]));
}
void orc_brdf_eval_n(const float n[3], const float wo[3], const float* wi, uint32_t count, const float base[3],
                     float metallic, float roughness, float* out) {
    for (uint32_t i = 0; i < count; ++i)
        vst(out + 3 * i, brdf_eval(vld(n), vld(wo), vld(wi + 3 * i), vld(base), metallic, roughness));
}
void orc_brdf_pdf_n(const float n[3], const float wo[3], const float* wi, uint32_t count, float roughness,
                    float* out) {
    for (uint32_t i = 0; i < count; ++i) out[i] = brdf_pdf(vld(wo), vld(wi + 3 * i), vld(n), roughness);
}
void orc_brdf_sample_n(const float n[3], const float wo[3], const float base[3], float metallic, float roughness,
                       const float* u, uint32_t count, float* wi, float* pdf, float* f) {
    for (uint32_t i = 0; i < count; ++i) {
        v3 w;
        vst(f + 3 * i, brdf_sample(vld(wo), vld(n), vld(base), metallic, roughness, u[2 * i], u[2 * i + 1], &w,
                                   pdf + i));
        vst(wi + 3 * i, w);
    }
}

/* the sampling pdf of brdf_sample evaluated at an arbitrary wi (world space);
 * used only for the MIS weights of triangle lights */
static float brdf_sample_pdf(v3 wo, v3 wi, v3 n, v3 base, float metallic, float roughness) {
    float r = clampf(roughness, 0.04f, 1.0f);
    float p_spec = p_spec_of(n, wo, base, metallic);
    v3 h = vnorm(vadd(wo, wi));
    float NdotH = fmaxf(vdot(n, h), 1e-6f);
    float VdotH = fmaxf(vdot(wo, h), 1e-6f);
    float D = distribution_ggx(NdotH, r * r);
    float pdf_spec = D * NdotH / fmaxf(4.0f * VdotH, 1e-6f);
    float pdf_diff = fmaxf(vdot(n, wi), 0.0f) * K_INV_PI;
    return fmaxf(p_spec * pdf_spec + (1.0f - p_spec) * pdf_diff, 1e-6f);
}

/* ---- material dispatch: reference metallic-roughness or pure Lambert ---- */
typedef struct { v3 base; float metallic, roughness; uint32_t lambert; v3 emissive; } matp;
static matp load_mat(const orc_scene* s, uint32_t mi) {
    const orc_material* m = &s->mat[mi];
    matp r;
    r.base = vld(m->base_color); r.metallic = m->metallic; r.roughness = m->roughness;
    r.lambert = (m->flags & ORC_MAT_PURE_LAMBERT) ? 1u : 0u;
    r.emissive = vld(m->emissive);
    return r;
}
static v3 mat_eval(const matp* m, v3 n, v3 wo, v3 wi) {
    if (m->lambert) return vscale(m->base, K_INV_PI);
    return brdf_eval(n, wo, wi, m->base, m->metallic, m->roughness);
}
static v3 mat_sample(const matp* m, v3 wo, v3 n, float ux, float uy, v3* wi, float* pdf) {
    if (m->lambert) {
        float c[3]; v3 T, B; build_onb(n, &T, &B);
        orc_sample_cosine_hemisphere(ux, uy, c, pdf);
        *wi = to_world(vld(c), T, B, n);
        return vscale(m->base, K_INV_PI);
    }
    return brdf_sample(wo, n, m->base, m->metallic, m->roughness, ux, uy, wi, pdf);
}
static float mat_sample_pdf(const matp* m, v3 wo, v3 wi, v3 n) {
    if (m->lambert) return fmaxf(vdot(n, wi), 0.0f) * K_INV_PI;
    return brdf_sample_pdf(wo, wi, n, m->base, m->metallic, m->roughness);
}
static float mat_ref_pdf(const matp* m, v3 wo, v3 wi, v3 n) {
    if (m->lambert) return fmaxf(vdot(n, wi), 0.0f) * K_INV_PI;
    return brdf_pdf(wo, wi, n, m->roughness);
}

/* ========================================================================== */
/* triangle lights (new: the reference has only point/directional lights)    */
/* ========================================================================== */
typedef struct { v3 p; v3 nl; float area; v3 Le; } tri_sample;
static tri_sample tri_light_point(const orc_scene* s, uint32_t tl, float u1, float u2) {
    uint32_t t = s->tl_tri[tl];
    const float* tv = &s->tri[(size_t)t * 9];
    v3 a = vld(tv), b = vld(tv + 3), c = vld(tv + 6);
    float su = sqrtf(u1);
    float b0 = 1.0f - su;
    float b1 = u2 * su;
    float b2 = 1.0f - b0 - b1;
    tri_sample r;
    r.p = vadd(vadd(vscale(a, b0), vscale(b, b1)), vscale(c, b2));
    v3 cr = vcross(vsub(b, a), vsub(c, a));
    float l2 = vlength(cr);
    r.area = 0.5f * l2;
    r.nl = vscale(cr, 1.0f / l2);
    const orc_mesh* me = &s->m[s->inst[s->tri_inst[t]].mesh_index];
    r.Le = vld(s->mat[me->material_index].emissive);
    return r;
}
/* 16-bit quantisation of the two triangle-sample numbers kept in a reservoir */
static uint32_t quant_uv(float u1, float u2) {
    uint32_t a = (uint32_t)(u1 * 65535.0f + 0.5f), b = (uint32_t)(u2 * 65535.0f + 0.5f);
    return a | (b << 16);
}
static void dequant_uv(uint32_t q, float* u1, float* u2) {
    *u1 = (float)(q & 0xffffu) * (1.0f / 65535.0f);
    *u2 = (float)(q >> 16) * (1.0f / 65535.0f);
}
static float tri_light_pdf(const orc_scene* s, uint32_t tri, v3 dir, float t) {
    /* solid-angle pdf of NEE for hitting triangle `tri` along unit `dir` at distance t */
    const float* tv = &s->tri[(size_t)tri * 9];
    v3 a = vld(tv), b = vld(tv + 3), c = vld(tv + 6);
    v3 cr = vcross(vsub(b, a), vsub(c, a));
    float l2 = vlength(cr);
    float area = 0.5f * l2;
    float cosl = fabsf(vdot(vscale(cr, 1.0f / l2), dir));
    float N = (float)(s->nl + s->ntl);
    return (t * t) / (cosl * area) / N;
}

/* ========================================================================== */
/* NEE -- pathtracer.slang:124-177 (+ triangle lights with power-heuristic MIS) */
/* ========================================================================== */
static v3 sample_lights(const orc_scene* s, v3 hit, v3 n, v3 wo, const matp* m, uint32_t* seed,
                        uint64_t* shadow_rays) {
    uint32_t N = s->nl + s->ntl;
    if (N == 0) return V(0, 0, 0);
    uint32_t li = (uint32_t)(orc_rand_float(seed) * (float)N) % N;
    v3 zero = V(0, 0, 0);
    if (li < s->nl) {
        const orc_light* L = &s->l[li];
        v3 ldir; float dist, pl;
        if (L->type == 0) {
            v3 to = vsub(vld(L->position), hit);
            dist = vlength(to);
            ldir = vdivs(to, dist);
            pl = dist * dist;
        } else {
            ldir = vnorm(vneg(vld(L->direction)));
            dist = 1e9f;
            pl = 1.0f;
        }
        if (vdot(ldir, n) <= 0.0f) return zero;
        (*shadow_rays)++;
        if (occluded_bvh(s, vadd(hit, vscale(n, 1e-4f)), ldir, 0.001f, dist - 0.01f)) return zero;
        v3 f = mat_eval(m, n, wo, ldir);
        float ndl = fmaxf(vdot(n, ldir), 0.0f);
        v3 E = vscale(vld(L->emission), L->intensity);
        float den = (L->type == 0) ? (pl + 1e-6f) : 1.0f;
        v3 irr = vdivs(E, den);
        float pb = mat_ref_pdf(m, wo, ldir, n);
        float mis = (L->type == 0) ? 1.0f : (pl * pl) / (pl * pl + pb * pb + 1e-10f);
        return vscale(vscale(vscale(vmul(f, irr), ndl), mis), (float)N);
    }
    float u1 = orc_rand_float(seed);
    float u2 = orc_rand_float(seed);
    tri_sample ts = tri_light_point(s, li - s->nl, u1, u2);
    v3 to = vsub(ts.p, hit);
    float d2 = vdot(to, to);
    float dist = sqrtf(d2);
    v3 l = vscale(to, 1.0f / dist);
    if (vdot(l, n) <= 0.0f) return zero;
    float cosl = fabsf(vdot(ts.nl, l));
    if (!(cosl > 0.0f)) return zero;
    (*shadow_rays)++;
    if (occluded_bvh(s, vadd(hit, vscale(n, 1e-4f)), l, 0.001f, dist - 0.01f)) return zero;
    v3 f = mat_eval(m, n, wo, l);
    float ndl = fmaxf(vdot(n, l), 0.0f);
    float pl = d2 / (cosl * ts.area) / (float)N;
    float pb = mat_sample_pdf(m, wo, l, n);
    float mis = (pl * pl) / (pl * pl + pb * pb);
    return vscale(vscale(vmul(f, ts.Le), ndl), mis / pl);
}

/* ========================================================================== */
/* path tracer -- pathtracer.slang:181-265                                   */
/* ========================================================================== */
static v3 pt_pixel(const orc_scene* s, const orc_camera* cam, const orc_pt_params* p, uint32_t px,
                   uint32_t py, orc_ray_counts* rc) {
    uint32_t seed = orc_make_seed(px, py, p->frame_seed);
    float o[3], d[3];
    orc_primary_ray(cam, px, py, &seed, o, d);
    v3 ro = vld(o), rd = vld(d);
    v3 L = V(0, 0, 0), thp = V(1, 1, 1);
    float last_pdf = 0.0f;
    for (uint32_t bounce = 0; bounce < p->max_bounces; ++bounce) {
        if (bounce == 0) rc->primary++; else rc->extension++;
        hitinfo h = trace(s, ro, rd, 0.001f, 1e9f);
        if (!h.hit) {
            L = vadd(L, vmul(thp, sky(rd)));
            break;
        }
        matp m = load_mat(s, h.mat);
        if (m.emissive.x > 0.0f || m.emissive.y > 0.0f || m.emissive.z > 0.0f) {
            uint32_t lo = s->tri_light_of[h.prim];
            if (bounce > 0 && lo != ~0u) {
                float pl = tri_light_pdf(s, h.prim, rd, h.t);
                float w = (last_pdf * last_pdf) / (last_pdf * last_pdf + pl * pl);
                L = vadd(L, vmul(thp, vscale(m.emissive, w)));
            } else {
                L = vadd(L, vmul(thp, m.emissive));
            }
            break;
        }
        v3 wo = vneg(rd);
        v3 n = h.normal;
        if (vdot(n, wo) < 0.0f) n = vneg(n);
        uint64_t sr = 0;
        v3 nee = sample_lights(s, h.pos, n, wo, &m, &seed, &sr);
        rc->shadow += sr;
        L = vadd(L, vmul(thp, nee));
        float ux = orc_rand_float(&seed);
        float uy = orc_rand_float(&seed);
        v3 wi; float pdf;
        v3 f = mat_sample(&m, wo, n, ux, uy, &wi, &pdf);
        if (pdf < 1e-6f) break;
        float ndw = fmaxf(vdot(n, wi), 0.0f);
        thp = vmul(thp, vdivs(vscale(f, ndw), pdf));
        if (bounce >= p->rr_start) {
            float survive = fminf(fmaxf(thp.x, fmaxf(thp.y, thp.z)), 0.95f);
            if (orc_rand_float(&seed) > survive) break;
            thp = vdivs(thp, survive);
        }
        ro = vadd(h.pos, vscale(n, 1e-4f));
        rd = wi;
        last_pdf = pdf;
    }
    return L;
}

void orc_pt_render(const orc_scene* s, const orc_camera* cam, const orc_pt_params* p,
                   const uint32_t* pixels, uint32_t npix, float* rgba, orc_ray_counts* counts,
                   int threads) {
    uint64_t cp = 0, ce = 0, cs = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : cp, ce, cs)
    for (int64_t i = 0; i < (int64_t)npix; ++i) {
        uint32_t pid = pixels[i];
        orc_ray_counts rc = {0, 0, 0};
        v3 L = pt_pixel(s, cam, p, pid % cam->screen_width, pid / cam->screen_width, &rc);
        rgba[i * 4 + 0] = L.x; rgba[i * 4 + 1] = L.y; rgba[i * 4 + 2] = L.z; rgba[i * 4 + 3] = 1.0f;
        cp += rc.primary; ce += rc.extension; cs += rc.shadow;
    }
    if (counts) { counts->primary = cp; counts->extension = ce; counts->shadow = cs; }
}

/* ========================================================================== */
/* accumulate -- accumulate.slang:21-46                                      */
/* ========================================================================== */
void orc_accumulate(const float* radiance, float* accum, size_t npix, uint32_t camera_moved,
                    uint32_t accumulated_spp) {
    for (size_t i = 0; i < npix * 4; ++i) {
        if (camera_moved != 0 || accumulated_spp == 0) { accum[i] = radiance[i]; continue; }
        float spp = (float)accumulated_spp;
        float weight = 1.0f / (spp + 1.0f);
        accum[i] = accum[i] * (spp / (spp + 1.0f)) + radiance[i] * weight;
    }
}
/* FP64 variant: running sum (mean = sum / spp after the call) */
void orc_accumulate_f64(const float* radiance, double* accum, size_t npix, uint32_t camera_moved,
                        uint32_t accumulated_spp) {
    for (size_t i = 0; i < npix * 4; ++i) {
        if (camera_moved != 0 || accumulated_spp == 0) accum[i] = (double)radiance[i];
        else accum[i] += (double)radiance[i];
    }
}

/* ========================================================================== */
/* ReSTIR DI -- restir_di.slang:57-472                                       */
/* ========================================================================== */
#define EMPTY_LIGHT 0xffffffffu

static float lum709(v3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

/* restir_di.slang:152-178; triangle lights add the cos_l / d^2 geometry term */
static float p_hat(const orc_scene* s, uint32_t li, uint32_t uvq, v3 hp, v3 n) {
    uint32_t N = s->nl + s->ntl;
    if (li >= N) return 0.0f;
    v3 l, irr;
    if (li < s->nl) {
        const orc_light* L = &s->l[li];
        if (L->type == 0) {
            v3 to = vsub(vld(L->position), hp);
            float d2 = fmaxf(vdot(to, to), 1e-6f);
            l = vnorm(to);
            irr = vdivs(vscale(vld(L->emission), L->intensity), d2);
        } else {
            l = vnorm(vneg(vld(L->direction)));
            irr = vscale(vld(L->emission), L->intensity);
        }
    } else {
        float u1, u2; dequant_uv(uvq, &u1, &u2);
        tri_sample ts = tri_light_point(s, li - s->nl, u1, u2);
        v3 to = vsub(ts.p, hp);
        float d2 = fmaxf(vdot(to, to), 1e-6f);
        l = vnorm(to);
        float cosl = fabsf(vdot(ts.nl, l));
        irr = vscale(ts.Le, cosl / d2);
    }
    float ndl = fmaxf(vdot(n, l), 0.0f);
    return ndl * lum709(irr);
}
static float src_area(const orc_scene* s, uint32_t li) {
    if (li < s->nl) return 1.0f;
    uint32_t t = s->tl_tri[li - s->nl];
    const float* tv = &s->tri[(size_t)t * 9];
    v3 a = vld(tv), b = vld(tv + 3), c = vld(tv + 6);
    return 0.5f * vlength(vcross(vsub(b, a), vsub(c, a)));
}

static void pack_res(float* r, uint32_t sel, float wsum, float W, uint32_t M) {
    r[0] = asfloat(sel); r[1] = wsum; r[2] = W; r[3] = asfloat(M);
}

void orc_di_initial(const orc_scene* s, const orc_camera* cam, const orc_di_params* p,
                    float* gpos, float* gnrm, float* res, uint32_t* res_uv, int threads) {
    uint32_t W = cam->screen_width, H = cam->screen_height;
    uint32_t N = s->nl + s->ntl;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)W * H; ++i) {
        uint32_t px = (uint32_t)(i % W), py = (uint32_t)(i / W);
        uint32_t seed = orc_make_seed(px, py, p->frame_seed);
        float o[3], d[3];
        orc_primary_ray(cam, px, py, &seed, o, d);
        hitinfo h = trace(s, vld(o), vld(d), 0.001f, 1e9f);
        float* gp = gpos + i * 4; float* gn = gnrm + i * 4; float* rs = res + i * 4;
        res_uv[i] = 0;
        if (!h.hit) {
            gp[0] = gp[1] = gp[2] = 0.0f; gp[3] = -1.0f;
            gn[0] = gn[1] = gn[2] = gn[3] = 0.0f;
            pack_res(rs, EMPTY_LIGHT, 0.0f, 0.0f, 0);
            continue;
        }
        v3 n = h.normal;
        if (vdot(n, vneg(vld(d))) < 0.0f) n = vneg(n);
        gp[0] = h.pos.x; gp[1] = h.pos.y; gp[2] = h.pos.z; gp[3] = h.t;
        gn[0] = n.x; gn[1] = n.y; gn[2] = n.z; gn[3] = asfloat(h.mat);
        if (N == 0) { pack_res(rs, EMPTY_LIGHT, 0.0f, 0.0f, 0); continue; }
        uint32_t sel = EMPTY_LIGHT, sel_uv = 0, Mc = 0;
        float wsum = 0.0f;
        for (uint32_t k = 0; k < p->num_initial_candidates; ++k) {
            uint32_t cand = (uint32_t)(orc_rand_float(&seed) * (float)N) % N;
            uint32_t q = 0;
            if (cand >= s->nl) {
                float a = orc_rand_float(&seed);
                float b = orc_rand_float(&seed);
                q = quant_uv(a, b);
            }
            float ph = p_hat(s, cand, q, h.pos, n);
            float wi = ph * (float)N * src_area(s, cand);
            wsum += wi;
            Mc++;
            if (wsum > 0.0f && orc_rand_float(&seed) < wi / wsum) { sel = cand; sel_uv = q; }
        }
        float Wf = 0.0f;
        if (sel != EMPTY_LIGHT) {
            float ph = p_hat(s, sel, sel_uv, h.pos, n);
            Wf = (ph > 1e-10f) ? (wsum / ((float)Mc * ph)) : 0.0f;
        }
        pack_res(rs, sel, wsum, Wf, Mc);
        res_uv[i] = sel_uv;
    }
}

void orc_di_temporal(const orc_scene* s, const orc_camera* cam, const orc_di_params* p,
                     const float* gpos, const float* gnrm, float* rc, uint32_t* rc_uv,
                     const float* rp, const uint32_t* rp_uv, int threads) {
    if (!p->temporal_enable) return;
    uint32_t W = cam->screen_width, H = cam->screen_height;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)W * H; ++i) {
        const float* gp = gpos + i * 4;
        if (gp[3] < 0.0f) continue;
        v3 hp = vld(gp), n = vld(gnrm + i * 4);
        float* cr = rc + i * 4;
        uint32_t csel = asuint(cr[0]), cM = asuint(cr[3]), cuv = rc_uv[i];
        float cw = cr[1];
        const float* pr = rp + i * 4;
        uint32_t psel = asuint(pr[0]), pM = asuint(pr[3]), puv = rp_uv[i];
        float pW = pr[2];
        if (psel == EMPTY_LIGHT || pM == 0) continue;
        uint32_t cap = p->num_initial_candidates * 20u;
        cap = cap < 1u ? 1u : (cap > 8192u ? 8192u : cap);
        pM = pM < cap ? pM : cap;
        float php = p_hat(s, psel, puv, hp, n);
        float contrib = php * pW * (float)pM;
        cw += contrib;
        uint32_t nM = cM + pM;
        if (nM > cap) { float sc = (float)cap / (float)nM; cw *= sc; cM = cap; }
        else cM = nM;
        uint32_t seed = orc_make_seed((uint32_t)(i % W), (uint32_t)(i / W), p->frame_seed + 1u);
        if (contrib > 0.0f && orc_rand_float(&seed) < contrib / cw) { csel = psel; cuv = puv; }
        float Wn = 0.0f;
        if (csel != EMPTY_LIGHT) {
            float ph = p_hat(s, csel, cuv, hp, n);
            Wn = (ph > 1e-10f) ? (cw / ((float)cM * ph)) : 0.0f;
            Wn = fminf(Wn, 1e4f);
        }
        pack_res(cr, csel, cw, Wn, cM);
        rc_uv[i] = cuv;
    }
}

static uint64_t g_di_shadow_rays = 0;
uint64_t orc_di_last_shadow_rays(void) { return g_di_shadow_rays; }

void orc_di_spatial_shade(const orc_scene* s, const orc_camera* cam, const orc_di_params* p,
                          const float* gpos, const float* gnrm, const float* res,
                          const uint32_t* res_uv, float* out, int threads) {
    uint32_t W = cam->screen_width, H = cam->screen_height;
    uint32_t N = s->nl + s->ntl;
    uint64_t shadow_rays = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : shadow_rays)
    for (int64_t i = 0; i < (int64_t)W * H; ++i) {
        uint32_t px = (uint32_t)(i % W), py = (uint32_t)(i / W);
        float* o = out + i * 4;
        const float* gp = gpos + i * 4;
        if (gp[3] < 0.0f) {
            v3 sk = sky(center_ray_dir(cam, px, py));
            o[0] = sk.x; o[1] = sk.y; o[2] = sk.z; o[3] = 1.0f;
            continue;
        }
        v3 hp = vld(gp);
        const float* gn = gnrm + i * 4;
        v3 n = vld(gn);
        matp m = load_mat(s, asuint(gn[3]));
        if (m.emissive.x > 0.0f || m.emissive.y > 0.0f || m.emissive.z > 0.0f) {
            o[0] = m.emissive.x; o[1] = m.emissive.y; o[2] = m.emissive.z; o[3] = 1.0f;
            continue;
        }
        const float* rr = res + i * 4;
        uint32_t rsel = asuint(rr[0]), rM = asuint(rr[3]), ruv = res_uv[i];
        float rw = rr[1], rW = rr[2];
        if (p->spatial_enable && p->spatial_num_neighbors > 0 && N > 0) {
            uint32_t seed = orc_make_seed(px, py, p->frame_seed + 2u);
            for (uint32_t k = 0; k < p->spatial_num_neighbors; ++k) {
                float angle = orc_rand_float(&seed) * 6.28318530717958f;
                float radius = orc_rand_float(&seed) * (float)p->spatial_radius;
                float sa, ca; orc_sincosf(angle, &sa, &ca);
                int ox = (int)(radius * ca), oy = (int)(radius * sa);
                int nx = (int)px + ox, ny = (int)py + oy;
                nx = nx < 0 ? 0 : (nx > (int)W - 1 ? (int)W - 1 : nx);
                ny = ny < 0 ? 0 : (ny > (int)H - 1 ? (int)H - 1 : ny);
                size_t j = (size_t)ny * W + (size_t)nx;
                if (gpos[j * 4 + 3] < 0.0f) continue;
                const float* nr = res + j * 4;
                uint32_t nsel = asuint(nr[0]), nM = asuint(nr[3]), nuv = res_uv[j];
                float nW = nr[2];
                if (nsel == EMPTY_LIGHT || nM == 0) continue;
                float ph = p_hat(s, nsel, nuv, hp, n);
                float contrib = ph * nW * (float)nM;
                rw += contrib;
                rM += nM;
                if (contrib > 0.0f && orc_rand_float(&seed) < contrib / rw) { rsel = nsel; ruv = nuv; }
            }
            if (rsel != EMPTY_LIGHT) {
                float ph = p_hat(s, rsel, ruv, hp, n);
                rW = (ph > 1e-10f) ? (rw / ((float)rM * ph)) : 0.0f;
            }
        }
        v3 outc = V(0, 0, 0);
        if (rsel != EMPTY_LIGHT && rW > 0.0f && N > 0) {
            v3 l, irr; float dist;
            if (rsel < s->nl) {
                const orc_light* L = &s->l[rsel];
                if (L->type == 0) {
                    v3 to = vsub(vld(L->position), hp);
                    dist = vlength(to);
                    l = vdivs(to, fmaxf(dist, 1e-6f));
                    irr = vdivs(vscale(vld(L->emission), L->intensity), fmaxf(dist * dist, 1e-6f));
                } else {
                    l = vnorm(vneg(vld(L->direction)));
                    dist = 1e9f;
                    irr = vscale(vld(L->emission), L->intensity);
                }
            } else {
                float u1, u2; dequant_uv(ruv, &u1, &u2);
                tri_sample ts = tri_light_point(s, rsel - s->nl, u1, u2);
                v3 to = vsub(ts.p, hp);
                dist = vlength(to);
                l = vdivs(to, fmaxf(dist, 1e-6f));
                float cosl = fabsf(vdot(ts.nl, l));
                irr = vscale(ts.Le, cosl / fmaxf(dist * dist, 1e-6f));
            }
            float ndl = fmaxf(vdot(n, l), 0.0f);
            if (ndl > 0.0f) {
                shadow_rays++;
                if (!occluded_bvh(s, vadd(hp, vscale(n, 1e-4f)), l, 0.001f, dist - 0.01f)) {
                    v3 wo = vnorm(vsub(vld(cam->position), hp));
                    v3 f = mat_eval(&m, n, wo, l);
                    outc = vscale(vscale(vmul(f, irr), ndl), rW);
                }
            }
        }
        o[0] = outc.x; o[1] = outc.y; o[2] = outc.z; o[3] = 1.0f;
    }
    g_di_shadow_rays = shadow_rays;
}

/* ========================================================================== */
/* A-trous -- atrous.slang:21-106 (reference mode)                            */
/* ========================================================================== */
/* The three edge-stopping weights exp(-d / sigma^2) (atrous.slang:83-96) are evaluated with
 * orc_wexpf, the fused-multiply-add exponential of the config mode (relative error about 2e-7),
 * as the kernel evaluates them; the reference's own exp is the driver's approximation, so no
 * exponential reproduces its bits. */
static float kernel_weight(int o) { return (o == -2 || o == 2) ? 1.0f : ((o == -1 || o == 1) ? 4.0f : 6.0f); }
static float luma601(v3 c) { return c.x * 0.299f + c.y * 0.587f + c.z * 0.114f; }

void orc_atrous_ref(const orc_atrous_params* p, const float* in, const float* gpos,
                    const float* gnrm, float* out) {
    int W = (int)p->width, H = (int)p->height, st = (int)p->stride;
    float sps = fmaxf(p->sigma_position * p->sigma_position, 1.0e-6f);
    float sns = fmaxf(p->sigma_normal * p->sigma_normal, 1.0e-6f);
    float sls = fmaxf(p->sigma_luminance * p->sigma_luminance, 1.0e-6f);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            size_t i = (size_t)y * W + x;
            const float* cp = gpos + i * 4;
            const float* cr = in + i * 4;
            float* o = out + i * 4;
            if (cp[3] < 0.0f) { memcpy(o, cr, 16); continue; }
            v3 cn = vnorm(vld(gnrm + i * 4));
            float cl = luma601(vld(cr));
            v3 acc = V(0, 0, 0);
            float wsum = 0.0f;
            for (int ty = -2; ty <= 2; ++ty) {
                for (int tx = -2; tx <= 2; ++tx) {
                    int sx = x + tx * st, sy = y + ty * st;
                    sx = sx < 0 ? 0 : (sx > W - 1 ? W - 1 : sx);
                    sy = sy < 0 ? 0 : (sy > H - 1 ? H - 1 : sy);
                    size_t j = (size_t)sy * W + sx;
                    const float* sp = gpos + j * 4;
                    if (sp[3] < 0.0f) continue;
                    v3 sr = vld(in + j * 4);
                    float bw = (kernel_weight(tx) * kernel_weight(ty)) / 256.0f;
                    v3 pd = vsub(vld(sp), vld(cp));
                    float pw = orc_wexpf(-vdot(pd, pd) / sps);
                    v3 sn = vnorm(vld(gnrm + j * 4));
                    float nd = fmaxf(0.0f, 1.0f - vdot(cn, sn));
                    float nw = orc_wexpf(-nd / sns);
                    float ld = luma601(sr) - cl;
                    float lw = orc_wexpf(-(ld * ld) / sls);
                    float w = bw * pw * nw * lw;
                    acc = vadd(acc, vscale(sr, w));
                    wsum += w;
                }
            }
            v3 r = (wsum > 0.0f) ? vdivs(acc, wsum) : vld(cr);
            o[0] = r.x; o[1] = r.y; o[2] = r.z; o[3] = cr[3];
        }
    }
}

/* ---- fp16 (IEEE binary16, round to nearest even) ---- */
uint16_t orc_f32_to_f16(float f) {
    uint32_t x = asuint(f);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u));
    if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); /* rounds to inf */
    if (ax < 0x38800000u) {                                    /* subnormal half */
        if (ax < 0x33000000u) return (uint16_t)sign;           /* < 2^-25 -> 0 (ties-even at 2^-25 handled below) */
        uint32_t e = ax >> 23;
        uint32_t m = (ax & 0x7fffffu) | 0x800000u;
        uint32_t shift = 126u - e;                             /* 14..24 */
        uint32_t hm = m >> shift;
        uint32_t rem = m & ((1u << shift) - 1u);
        uint32_t half = 1u << (shift - 1u);
        if (rem > half || (rem == half && (hm & 1u))) hm++;
        return (uint16_t)(sign | hm);
    }
    uint32_t hx = ((ax - 0x38000000u) >> 13);
    uint32_t rem = ax & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (hx & 1u))) hx++;
    return (uint16_t)(sign | hx);
}
float orc_f16_to_f32(uint16_t h) {
    uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    if (e == 0) {
        float v = (float)m * (1.0f / 16777216.0f);
        return (h & 0x8000u) ? -v : v;
    }
    if (e == 31) return asfloat(sign | 0x7f800000u | (m << 13));
    return asfloat(sign | ((e + 112u) << 23) | (m << 13));
}

/* config-mode inputs from a rendered frame (the library's rrh_atrous_prepare; no reference
 * counterpart: the reference filters RGBA32F radiance with its fp32 G-buffer): fp16 radiance
 * with the population variance of the luminance (Rec.709) over the valid pixels of the
 * clamped 3x3 window in alpha, and fp16 (normal, linear depth), the depth being the hit
 * point's distance along the camera's forward direction (0 on a miss). */
void orc_atrous_prepare(uint32_t W, uint32_t H, const float cpos[3], const float cdir[3], const float* rad,
                        const float* gpos, const float* gnrm, uint16_t* rad16, uint16_t* gb16) {
    v3 cp = vld(cpos), cd = vld(cdir);
#pragma omp parallel for schedule(static)
    for (int64_t y = 0; y < (int64_t)H; ++y) {
        for (int64_t x = 0; x < (int64_t)W; ++x) {
            size_t i = (size_t)y * W + (size_t)x;
            float m1 = 0.0f, m2 = 0.0f, n = 0.0f;
            for (int dy = -1; dy <= 1; ++dy) {
                for (int dx = -1; dx <= 1; ++dx) {
                    int64_t sx = x + dx, sy = y + dy;
                    sx = sx < 0 ? 0 : (sx > (int64_t)W - 1 ? (int64_t)W - 1 : sx);
                    sy = sy < 0 ? 0 : (sy > (int64_t)H - 1 ? (int64_t)H - 1 : sy);
                    size_t j = (size_t)sy * W + (size_t)sx;
                    if (gpos[j * 4 + 3] < 0.0f) continue;
                    float l = lum709(vld(rad + j * 4));
                    m1 += l; m2 += l * l; n += 1.0f;
                }
            }
            float var = 0.0f;
            if (n > 0.0f) { m1 = m1 / n; m2 = m2 / n; var = fmaxf(0.0f, m2 - m1 * m1); }
            const float* c = rad + i * 4;
            rad16[i * 4] = orc_f32_to_f16(c[0]); rad16[i * 4 + 1] = orc_f32_to_f16(c[1]);
            rad16[i * 4 + 2] = orc_f32_to_f16(c[2]); rad16[i * 4 + 3] = orc_f32_to_f16(var);
            const float* gp = gpos + i * 4; const float* gn = gnrm + i * 4;
            float z = (gp[3] < 0.0f) ? 0.0f : vdot(vsub(vld(gp), cp), cd);
            gb16[i * 4] = orc_f32_to_f16(gn[0]); gb16[i * 4 + 1] = orc_f32_to_f16(gn[1]);
            gb16[i * 4 + 2] = orc_f32_to_f16(gn[2]); gb16[i * 4 + 3] = orc_f32_to_f16(z);
        }
    }
}

/* config mode: SVGF edge-stopping (Schied et al. 2017, eqs. 3-5) on fp16
 * buffers.  in.a carries the luminance variance; gbuf = (n.xyz, linear depth).
 * The weights' arithmetic is this build's own (the reference has no config mode), chosen for
 * few operations per tap: max(0, n.n)^sigma_n by repeated squaring when sigma_n is 2^k
 * (k <= 10; otherwise exp(sigma_n log x)); one exp per tap (orc_wexpf); the normal dot product,
 * the exp's argument and the weighted sums as fused multiply-adds; the luminance term |dl| x (1 / phi_l) with the
 * reciprocal per pixel; the depth term |dz| x (1 / (|gx tx + gy ty| + 1e-3)) with
 * gx = sigma_z dz/dx S, gy = sigma_z dz/dy S per pixel and the 25 reciprocals per pixel (taps
 * o and -o share one value exactly). */
static float powi_pos(float x, float e) { return (x > 0.0f) ? orc_expf(e * orc_logf(x)) : 0.0f; }
/* exp of the config-mode weights (this build's own definition, as on the GPU): x = k ln2 + r by
 * a two-part fused reduction, e^r by a degree-6 Taylor polynomial in fmaf (each step rounds once,
 * so the kernel gives the same bits), times 2^k; relative error about 2e-7 for |r| <= ln2 / 2 */
float orc_wexpf(float x) {
    if (x < -87.0f) return 0.0f;
    if (x > 88.0f) x = 88.0f;
    const float k = floorf(fmaf(x, 1.44269502f, 0.5f));
    float r = fmaf(-k, 0.693145751953125f, x);
    r = fmaf(-k, 1.42860677e-6f, r);
    float q = fmaf(r, 1.38888892e-3f, 8.33333377e-3f);
    q = fmaf(r, q, 4.16666679e-2f);
    q = fmaf(r, q, 0.166666672f);
    q = fmaf(r, q, 0.5f);
    q = fmaf(r, q, 1.0f);
    q = fmaf(r, q, 1.0f);
    uint32_t bits = (uint32_t)((int)k + 127) << 23;
    float sc;
    memcpy(&sc, &bits, 4);
    return q * sc;
}
static int pow2_exponent(float e) {
    for (int k = 0; k <= 10; ++k)
        if (e == (float)(1 << k)) return k;
    return -1;
}

void orc_atrous_cfg(const orc_atrous_params* p, const uint16_t* in, const uint16_t* gbuf,
                    uint16_t* out) {
    int W = (int)p->width, H = (int)p->height, st = (int)p->stride;
    const int nsq = pow2_exponent(p->sigma_normal);
    const float gk[3] = {0.25f, 0.5f, 0.25f};
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            size_t i = (size_t)y * W + x;
            const uint16_t* cg = gbuf + i * 4;
            float zp = orc_f16_to_f32(cg[3]);
            if (!(zp > 0.0f)) { memcpy(out + i * 4, in + i * 4, 8); continue; }
            v3 np = V(orc_f16_to_f32(cg[0]), orc_f16_to_f32(cg[1]), orc_f16_to_f32(cg[2]));
            v3 cp = V(orc_f16_to_f32(in[i * 4]), orc_f16_to_f32(in[i * 4 + 1]), orc_f16_to_f32(in[i * 4 + 2]));
            float lp = lum709(cp);
            /* 3x3 Gaussian of the variance */
            float var = 0.0f;
            for (int dy = -1; dy <= 1; ++dy) {
                for (int dx = -1; dx <= 1; ++dx) {
                    int sx = x + dx, sy = y + dy;
                    sx = sx < 0 ? 0 : (sx > W - 1 ? W - 1 : sx);
                    sy = sy < 0 ? 0 : (sy > H - 1 ? H - 1 : sy);
                    var += gk[dx + 1] * gk[dy + 1] * orc_f16_to_f32(in[((size_t)sy * W + sx) * 4 + 3]);
                }
            }
            float phil = p->sigma_luminance * sqrtf(fmaxf(0.0f, var)) + 1e-10f;
            float inv_phil = 1.0f / phil;
            /* screen-space depth gradient by central differences */
            int xm = x > 0 ? x - 1 : 0, xp = x < W - 1 ? x + 1 : W - 1;
            int ym = y > 0 ? y - 1 : 0, yp = y < H - 1 ? y + 1 : H - 1;
            float dzdx = (orc_f16_to_f32(gbuf[((size_t)y * W + xp) * 4 + 3]) - orc_f16_to_f32(gbuf[((size_t)y * W + xm) * 4 + 3])) * 0.5f;
            float dzdy = (orc_f16_to_f32(gbuf[((size_t)yp * W + x) * 4 + 3]) - orc_f16_to_f32(gbuf[((size_t)ym * W + x) * 4 + 3])) * 0.5f;
            float gx = p->sigma_position * dzdx * (float)st, gy = p->sigma_position * dzdy * (float)st;
            float inv_den[5][5];
            for (int ty = -2; ty <= 2; ++ty)
                for (int tx = -2; tx <= 2; ++tx)
                    inv_den[ty + 2][tx + 2] = 1.0f / (fabsf(gx * (float)tx + gy * (float)ty) + 1e-3f);
            v3 acc = V(0, 0, 0);
            float accv = 0.0f, wsum = 0.0f;
            for (int ty = -2; ty <= 2; ++ty) {
                for (int tx = -2; tx <= 2; ++tx) {
                    int sx = x + tx * st, sy = y + ty * st;
                    sx = sx < 0 ? 0 : (sx > W - 1 ? W - 1 : sx);
                    sy = sy < 0 ? 0 : (sy > H - 1 ? H - 1 : sy);
                    size_t j = (size_t)sy * W + sx;
                    const uint16_t* sg = gbuf + j * 4;
                    float zq = orc_f16_to_f32(sg[3]);
                    if (!(zq > 0.0f)) continue;
                    v3 nq = V(orc_f16_to_f32(sg[0]), orc_f16_to_f32(sg[1]), orc_f16_to_f32(sg[2]));
                    v3 cq = V(orc_f16_to_f32(in[j * 4]), orc_f16_to_f32(in[j * 4 + 1]), orc_f16_to_f32(in[j * 4 + 2]));
                    float vq = orc_f16_to_f32(in[j * 4 + 3]);
                    float bw = (kernel_weight(tx) * kernel_weight(ty)) / 256.0f;
                    float wz = fabsf(zp - zq) * inv_den[ty + 2][tx + 2];
                    float wn = fmaxf(0.0f, fmaf(np.z, nq.z, fmaf(np.y, nq.y, np.x * nq.x)));
                    if (nsq >= 0) { for (int k = 0; k < nsq; ++k) wn = wn * wn; }
                    else wn = powi_pos(wn, p->sigma_normal);
                    float w = bw * orc_wexpf(fmaf(-fabsf(lp - lum709(cq)), inv_phil, -wz)) * wn;
                    acc.x = fmaf(cq.x, w, acc.x);
                    acc.y = fmaf(cq.y, w, acc.y);
                    acc.z = fmaf(cq.z, w, acc.z);
                    accv = fmaf(vq, w * w, accv);
                    wsum += w;
                }
            }
            uint16_t* o = out + i * 4;
            if (wsum > 0.0f) {
                o[0] = orc_f32_to_f16(acc.x / wsum); o[1] = orc_f32_to_f16(acc.y / wsum);
                o[2] = orc_f32_to_f16(acc.z / wsum); o[3] = orc_f32_to_f16(accv / (wsum * wsum));
            } else {
                memcpy(o, in + i * 4, 8);
            }
        }
    }
}

/* ========================================================================== */
/* vMF guiding grid -- GuidingGrid.slang:54-189, restir_pt.slang:520-563      */
/* ========================================================================== */
uint32_t orc_guiding_cell_index(const float p[3], uint32_t bsdf_id, const float cam[3],
                                float base_cell, float slope, uint32_t cell_count) {
    v3 wp = vld(p);
    float vd = vlength(vsub(wp, vld(cam)));
    float cell = fmaxf(base_cell + slope * vd, 1.0e-4f);
    int qx = (int)floorf(wp.x / cell), qy = (int)floorf(wp.y / cell), qz = (int)floorf(wp.z / cell);
    uint32_t h = orc_pcg_hash((uint32_t)qx * 73856093u ^ (uint32_t)qy * 19349663u ^
                              (uint32_t)qz * 83492791u ^ bsdf_id * 2654435761u);
    return h % (cell_count > 1u ? cell_count : 1u);
}
void orc_guiding_splat(uint32_t* accum, uint32_t cell, const float dir[3]) {
    uint32_t* a = accum + (size_t)cell * 4;
    a[0] += (uint32_t)(int32_t)roundf(dir[0] * 1048576.0f);
    a[1] += (uint32_t)(int32_t)roundf(dir[1] * 1048576.0f);
    a[2] += (uint32_t)(int32_t)roundf(dir[2] * 1048576.0f);
    a[3] += 1u;
}
void orc_guiding_fit_cell(uint32_t* accum, float* fitted, uint32_t cell, uint32_t min_samples,
                          float kappa_max) {
    uint32_t* a = accum + (size_t)cell * 4;
    float* f = fitted + (size_t)cell * 4;
    int32_t ix = (int32_t)a[0], iy = (int32_t)a[1], iz = (int32_t)a[2];
    uint32_t n = a[3];
    a[0] = a[1] = a[2] = a[3] = 0;
    if (n < (min_samples > 1u ? min_samples : 1u)) { f[3] = -1.0f; return; }
    v3 sum = vscale(V((float)ix, (float)iy, (float)iz), 1.0f / 1048576.0f);
    float len = vlength(sum);
    float r = len / (float)n;
    r = clampf(r, 1.0e-4f, 0.9999f);
    v3 mu = (len > 1.0e-6f) ? vdivs(sum, len) : V(0.0f, 0.0f, 1.0f);
    float kappa = r * (3.0f - r * r) / (1.0f - r * r);
    kappa = clampf(kappa, 0.5f, kappa_max);
    f[0] = mu.x; f[1] = mu.y; f[2] = mu.z; f[3] = kappa;
}
static v3 vmf_sample(v3 mu, float kappa, float u0, float u1, float* pdf) {
    float k = fmaxf(kappa, 1.0e-3f);
    float w = 1.0f + (1.0f / k) * orc_logf(u0 + (1.0f - u0) * orc_expf(-2.0f * k));
    w = clampf(w, -1.0f, 1.0f);
    float s = sqrtf(fmaxf(0.0f, 1.0f - w * w));
    float phi = K_TWO_PI * u1;
    float sp, cp; orc_sincosf(phi, &sp, &cp);
    v3 local = V(s * cp, s * sp, w);
    v3 T, B; build_onb(mu, &T, &B);
    v3 dir = to_world(local, T, B, mu);
    float denom = K_TWO_PI * (1.0f - orc_expf(-2.0f * k));
    *pdf = k * orc_expf(k * (w - 1.0f)) / fmaxf(denom, 1.0e-8f);
    return dir;
}
void orc_vmf_sample(const float mu[3], float kappa, float u0, float u1, float dir[3], float* pdf) {
    vst(dir, vmf_sample(vld(mu), kappa, u0, u1, pdf));
}
static float vmf_pdf(v3 mu, float kappa, v3 dir) {
    float k = fmaxf(kappa, 1.0e-3f);
    float w = clampf(vdot(mu, dir), -1.0f, 1.0f);
    float denom = K_TWO_PI * (1.0f - orc_expf(-2.0f * k));
    return k * orc_expf(k * (w - 1.0f)) / fmaxf(denom, 1.0e-8f);
}
float orc_vmf_pdf(const float mu[3], float kappa, const float dir[3]) {
    return vmf_pdf(vld(mu), kappa, vld(dir));
}
void orc_sample_primary_dir(const float* fitted, const float pos[3], const float n[3],
                            uint32_t bsdf_id, const float cam[3], float base_cell, float slope,
                            uint32_t cell_count, uint32_t guiding_enable, float alpha_in,
                            uint32_t* seed, float wi_out[3], float* src_pdf) {
    float alpha = 0.0f;
    int valid = 0;
    v3 mu = V(0, 0, 0); float kappa = 0.0f;
    if (guiding_enable != 0u && alpha_in > 0.0f) {
        uint32_t c = orc_guiding_cell_index(pos, bsdf_id, cam, base_cell, slope, cell_count);
        const float* f = fitted + (size_t)c * 4;
        mu = vld(f); kappa = f[3];
        valid = (kappa > 0.0f) && (vdot(mu, mu) > 0.25f);
        if (valid) alpha = alpha_in;
    }
    v3 N = vld(n), T, B; build_onb(N, &T, &B);
    v3 wi;
    if (orc_rand_float(seed) < alpha) {
        float u0 = orc_rand_float(seed), u1 = orc_rand_float(seed), vp;
        wi = vmf_sample(mu, kappa, u0, u1, &vp);
    } else {
        float u0 = orc_rand_float(seed), u1 = orc_rand_float(seed), c[3], cp;
        orc_sample_cosine_hemisphere(u0, u1, c, &cp);
        wi = to_world(vld(c), T, B, N);
    }
    float ci = vdot(N, wi);
    float pc = (ci > 0.0f) ? (ci * K_INV_PI) : 0.0f;
    float pv = valid ? vmf_pdf(mu, kappa, wi) : 0.0f;
    *src_pdf = alpha * pv + (1.0f - alpha) * pc;
    vst(wi_out, wi);
}

/* ========================================================================== */
/* 8-lobe vMF mixture EM (FP64) -- new; the M-step for one lobe is the
 * closed form of GuidingGrid.slang:100-137 on weighted directions           */
/* ========================================================================== */
static double vmf_pdf_d(const float* lobe, const float d[3]) {
    double k = lobe[3] > 1e-3f ? lobe[3] : 1e-3;
    double w = (double)lobe[0] * d[0] + (double)lobe[1] * d[1] + (double)lobe[2] * d[2];
    w = w < -1.0 ? -1.0 : (w > 1.0 ? 1.0 : w);
    return k * exp(k * (w - 1.0)) / (2.0 * 3.14159265358979323846 * (1.0 - exp(-2.0 * k)));
}
double orc_vmm_pdf(const float* lobes_cell, uint32_t K, const float dir[3]) {
    double p = 0.0;
    for (uint32_t k = 0; k < K; ++k) p += (double)lobes_cell[k * 5 + 4] * vmf_pdf_d(lobes_cell + k * 5, dir);
    return p;
}
static uint32_t grid_cell(const float p[3], const float bmin[3], const float bmax[3], uint32_t res) {
    uint32_t c[3];
    for (int k = 0; k < 3; ++k) {
        float f = (p[k] - bmin[k]) / (bmax[k] - bmin[k]) * (float)res;
        int q = (int)floorf(f);
        q = q < 0 ? 0 : (q > (int)res - 1 ? (int)res - 1 : q);
        c[k] = (uint32_t)q;
    }
    return (c[2] * res + c[1]) * res + c[0];
}
void orc_vmm_estep(const float* lobes, uint32_t K, const float* pos, const float* dir,
                   const float* w, uint32_t n, const float bmin[3], const float bmax[3],
                   uint32_t res, double* stats) {
    /* stats per cell: K x (s0, s1x, s1y, s1z) */
    for (uint32_t i = 0; i < n; ++i) {
        if (!(w[i] > 0.0f)) continue;
        uint32_t c = grid_cell(pos + i * 3, bmin, bmax, res);
        const float* lc = lobes + (size_t)c * K * 5;
        double g[64], tot = 0.0;
        for (uint32_t k = 0; k < K; ++k) { g[k] = (double)lc[k * 5 + 4] * vmf_pdf_d(lc + k * 5, dir + i * 3); tot += g[k]; }
        double* st = stats + (size_t)c * K * 4;
        for (uint32_t k = 0; k < K; ++k) {
            double gk = tot > 0.0 ? g[k] / tot : 1.0 / K;
            double wg = (double)w[i] * gk;
            st[k * 4 + 0] += wg;
            st[k * 4 + 1] += wg * dir[i * 3 + 0];
            st[k * 4 + 2] += wg * dir[i * 3 + 1];
            st[k * 4 + 3] += wg * dir[i * 3 + 2];
        }
    }
}
void orc_vmm_mstep(const double* stats, uint32_t ncell, uint32_t K, float kappa_max,
                   double min_weight, float* lobes) {
    for (uint32_t c = 0; c < ncell; ++c) {
        const double* st = stats + (size_t)c * K * 4;
        double W = 0.0;
        for (uint32_t k = 0; k < K; ++k) W += st[k * 4];
        if (!(W > min_weight)) continue; /* keep the previous lobes */
        float* lc = lobes + (size_t)c * K * 5;
        for (uint32_t k = 0; k < K; ++k) {
            double s0 = st[k * 4];
            double sx = st[k * 4 + 1], sy = st[k * 4 + 2], sz = st[k * 4 + 3];
            double len = sqrt(sx * sx + sy * sy + sz * sz);
            if (!(s0 > 0.0)) { lc[k * 5 + 4] = 0.0f; continue; }
            double r = len / s0;
            r = r < 1e-4 ? 1e-4 : (r > 0.9999 ? 0.9999 : r);
            if (len > 1e-12) { lc[k * 5] = (float)(sx / len); lc[k * 5 + 1] = (float)(sy / len); lc[k * 5 + 2] = (float)(sz / len); }
            double kap = r * (3.0 - r * r) / (1.0 - r * r);
            kap = kap < 0.5 ? 0.5 : (kap > kappa_max ? kappa_max : kap);
            lc[k * 5 + 3] = (float)kap;
            lc[k * 5 + 4] = (float)(s0 / W);
        }
    }
}

/* ========================================================================== */
/* ReSTIR PT, reconnection shift -- restir_pt.slang:98-1001 (stages :569-771,
 * :775-872, :876-1001).  Rays are two-sided: the shaders ask for
 * RAY_FLAG_CULL_BACK_FACING_TRIANGLES (restir_pt.slang:104,111, restir_gi.slang:75,82),
 * but every TLAS instance carries VK_GEOMETRY_INSTANCE_TRIANGLE_FACING_CULL_DISABLE_BIT_KHR
 * (AccelStructure.cpp:90, Scene.cpp:490), which switches that culling off.
 * orc_set_restir_cull(1) restates what the ray flag alone would do.  1/sqrtf where
 * the shader writes rsqrt.  Scenes whose emissive triangles are lights: light sampling
 * covers them (area sampling, no MIS) and the reconnection vertex's own emission is
 * left to the direct lighting pass.                                             */
/* ========================================================================== */
static int g_restir_cull = 0;
void orc_set_restir_cull(int on) { g_restir_cull = on ? 1 : 0; }
typedef struct { v3 rc_pos, rc_n, suffix, prefix; float w_sum, W; uint32_t M; float valid; } rpt_res;

static rpt_res rpt_load(const float* res, size_t npix, size_t i) {
    const float* p = res + i * 4; const float* n = res + (npix + i) * 4;
    const float* r = res + (2 * npix + i) * 4; const float* e = res + (3 * npix + i) * 4;
    rpt_res s;
    s.rc_pos = vld(p); s.w_sum = p[3];
    s.rc_n = vld(n); s.M = asuint(n[3]);
    s.suffix = vld(r); s.W = r[3];
    s.prefix = vld(e); s.valid = e[3];
    return s;
}
static void rpt_store(float* res, size_t npix, size_t i, const rpt_res* s) {
    float* p = res + i * 4; float* n = res + (npix + i) * 4;
    float* r = res + (2 * npix + i) * 4; float* e = res + (3 * npix + i) * 4;
    vst(p, s->rc_pos); p[3] = s->w_sum;
    vst(n, s->rc_n); n[3] = asfloat(s->M);
    vst(r, s->suffix); r[3] = s->W;
    vst(e, s->prefix); e[3] = s->valid;
}
static void rpt_store_empty(float* res, size_t npix, size_t i) {
    for (int k = 0; k < 4; ++k) memset(res + (k * npix + i) * 4, 0, 16);
}
static int hit_cull(const orc_scene* s, v3 o, v3 d, float tmin, float tmax) {
    return trace_bvh_cull(s, o, d, tmin, tmax, g_restir_cull).prim != ~0u;
}

/* sample_lights, restir_pt.slang:171-222 */
static v3 rpt_lights(const orc_scene* s, v3 hit, v3 n, v3 wo, const matp* m, float bias, uint32_t* seed) {
    const v3 zero = V(0, 0, 0);
    uint32_t N = s->nl + s->ntl;
    if (N == 0) return zero;
    uint32_t li = (uint32_t)(orc_rand_float(seed) * (float)N) % N;
    if (li < s->nl) {
        const orc_light* L = &s->l[li];
        v3 ldir; float dist, pl;
        if (L->type == 0) {
            v3 to = vsub(vld(L->position), hit);
            dist = vlength(to);
            ldir = vdivs(to, fmaxf(dist, 1e-6f));
            pl = dist * dist;
        } else {
            ldir = vnorm(vneg(vld(L->direction)));
            dist = 1e9f;
            pl = 1.0f;
        }
        if (vdot(ldir, n) <= 0.0f) return zero;
        if (hit_cull(s, vadd(hit, vscale(n, bias)), ldir, bias, dist - 0.01f)) return zero;
        v3 f = mat_eval(m, n, wo, ldir);
        float ndl = fmaxf(vdot(n, ldir), 0.0f);
        v3 irr = vdivs(vscale(vld(L->emission), L->intensity), (L->type == 0) ? (pl + 1e-6f) : 1.0f);
        float pb = mat_ref_pdf(m, wo, ldir, n);
        float mis = (L->type == 0) ? 1.0f : (pl * pl) / (pl * pl + pb * pb + 1e-10f);
        return vscale(vscale(vscale(vmul(f, irr), ndl), mis), (float)N);
    }
    float u1 = orc_rand_float(seed);
    float u2 = orc_rand_float(seed);
    tri_sample ts = tri_light_point(s, li - s->nl, u1, u2);
    v3 to = vsub(ts.p, hit);
    float d2 = vdot(to, to);
    float dist = sqrtf(d2);
    v3 l = vscale(to, 1.0f / dist);
    if (vdot(l, n) <= 0.0f) return zero;
    float cosl = fabsf(vdot(ts.nl, l));
    if (!(cosl > 0.0f)) return zero;
    if (hit_cull(s, vadd(hit, vscale(n, bias)), l, bias, dist - 0.01f)) return zero;
    v3 f = mat_eval(m, n, wo, l);
    float ndl = fmaxf(vdot(n, l), 0.0f);
    float pl = d2 / (cosl * ts.area) / (float)N;
    return vscale(vscale(vmul(f, ts.Le), ndl), 1.0f / pl);
}

/* evaluate_contribution / target_function, restir_pt.slang:464-489 */
static v3 rpt_contribution(v3 pp, v3 pn, const matp* pm, v3 cam, const rpt_res* r) {
    v3 to = vsub(r->rc_pos, pp);
    float d2 = vdot(to, to);
    if (d2 <= 1e-8f) return V(0, 0, 0);
    v3 wi = vscale(to, 1.0f / sqrtf(d2));
    float cx = fmaxf(vdot(pn, wi), 0.0f);
    if (cx <= 0.0f) return V(0, 0, 0);
    v3 wo = vnorm(vsub(cam, pp));
    v3 f = mat_eval(pm, pn, wo, wi);
    return vmul(vmul(vscale(f, cx), r->prefix), r->suffix);
}
static float rpt_target(v3 pp, v3 pn, const matp* pm, v3 cam, const rpt_res* r) {
    return lum709(rpt_contribution(pp, pn, pm, cam, r));
}

/* reconnection_jacobian, restir_pt.slang:427-447 */
static float rpt_jacobian(v3 src, v3 dst, v3 rc, v3 rn) {
    v3 sd = vsub(rc, src), td = vsub(rc, dst);
    float s2 = vdot(sd, sd), t2 = vdot(td, td);
    if (s2 <= 1e-8f || t2 <= 1e-8f) return 0.0f;
    v3 sdir = vscale(sd, 1.0f / sqrtf(s2)), tdir = vscale(td, 1.0f / sqrtf(t2));
    float sc = fabsf(vdot(rn, vneg(sdir))), tc = fabsf(vdot(rn, vneg(tdir)));
    if (sc <= 0.05f || tc <= 0.05f) return 0.0f;
    return clampf((tc * s2) / (sc * t2), 0.0f, 3.0f);
}

/* trace_visibility, restir_pt.slang:449-458 */
static int rpt_visible(const orc_scene* s, v3 o, v3 n, v3 target, float bias) {
    v3 delta = vsub(target, o);
    float dist = vlength(delta);
    if (dist <= bias) return 0;
    return !hit_cull(s, vadd(o, vscale(n, bias)), vdivs(delta, dist), bias, dist - bias);
}

/* load_surface_from_images, restir_pt.slang:285-305 */
static int rpt_surface(const float* pos, const float* nrm, size_t i, v3* p, v3* n, uint32_t* mat) {
    const float* a = pos + i * 4;
    if (a[3] < 0.0f) return 0;
    const float* b = nrm + i * 4;
    *p = vld(a);
    *n = vnorm(vld(b));
    *mat = asuint(b[3]);
    return 1;
}

void orc_rpt_initial(const orc_scene* s, const orc_camera* cam, const orc_rpt_params* p, const float* fitted,
                     float base_cell, float slope, uint32_t cell_count, const float* gpos, const float* gnrm,
                     float* res, float* hpos, float* hnrm, int threads) {
    const uint32_t W = cam->screen_width, H = cam->screen_height;
    const size_t npix = (size_t)W * H;
    const v3 cp = vld(cam->position);
    const int rc_emission = s->ntl == 0;
    const float alpha = p->guiding_enable ? p->guiding_alpha : 0.0f;
    const uint32_t max_b = p->max_bounces > 1u ? p->max_bounces : 1u;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t ii = 0; ii < (int64_t)npix; ++ii) {
        const size_t i = (size_t)ii;
        const uint32_t px = (uint32_t)(i % W), py = (uint32_t)(i / W);
        v3 pp, pn; uint32_t pmi;
        if (!rpt_surface(gpos, gnrm, i, &pp, &pn, &pmi)) {
            float* h = hpos + i * 4; h[0] = h[1] = h[2] = 0.0f; h[3] = -1.0f;
            memset(hnrm + i * 4, 0, 16);
            rpt_store_empty(res, npix, i);
            continue;
        }
        memcpy(hpos + i * 4, gpos + i * 4, 16);
        memcpy(hnrm + i * 4, gnrm + i * 4, 16);
        const matp pm = load_mat(s, pmi);
        const float bias = p->ray_bias;
        uint32_t seed = orc_make_seed(px, py, p->frame_seed);
        rpt_res sel; memset(&sel, 0, sizeof(sel));
        float w_sum = 0.0f;
        uint32_t Mc = 0;
        for (uint32_t c = 0; c < p->num_initial_candidates; ++c) {
            float src_pdf, wa[3], ppa[3], pna[3], cpa[3];
            vst(ppa, pp); vst(pna, pn); vst(cpa, cp);
            orc_sample_primary_dir(fitted, ppa, pna, pmi, cpa, base_cell, slope, cell_count,
                                   fitted != NULL, alpha, &seed, wa, &src_pdf);
            const v3 wi0 = vld(wa);
            if (src_pdf <= 1e-6f || vdot(pn, wi0) <= 0.0f) continue;
            orc_hit h0 = trace_bvh_cull(s, vadd(pp, vscale(pn, bias)), wi0, bias, 1e9f, g_restir_cull);
            if (h0.prim == ~0u) continue;
            hitinfo rc = fetch_hit(s, h0);
            v3 rn = rc.normal;
            if (vdot(rn, wi0) > 0.0f) rn = vneg(rn);
            const matp rm = load_mat(s, rc.mat);
            v3 suffix = rc_emission ? rm.emissive : V(0, 0, 0);
            const v3 rwo = vneg(wi0);
            suffix = vadd(suffix, rpt_lights(s, rc.pos, rn, rwo, &rm, bias, &seed));
            v3 cur_p = rc.pos, cur_n = rn, cur_wo = rwo;
            matp cur_m = rm;
            v3 thp = V(1, 1, 1);
            for (uint32_t b = 1; b < max_b; ++b) {
                v3 wi; float pdf;
                float ux = orc_rand_float(&seed), uy = orc_rand_float(&seed);
                v3 f = mat_sample(&cur_m, cur_wo, cur_n, ux, uy, &wi, &pdf);
                if (pdf <= 1e-6f || vdot(cur_n, wi) <= 0.0f) break;
                float cosc = fmaxf(vdot(cur_n, wi), 0.0f);
                thp = vmul(thp, vdivs(vscale(f, cosc), pdf));
                if (fmaxf(thp.x, fmaxf(thp.y, thp.z)) <= 0.0f) break;
                orc_hit h = trace_bvh_cull(s, vadd(cur_p, vscale(cur_n, bias)), wi, bias, 1e9f, g_restir_cull);
                if (h.prim == ~0u) { suffix = vadd(suffix, vmul(thp, sky(wi))); break; }
                hitinfo hi = fetch_hit(s, h);
                v3 hn = hi.normal;
                if (vdot(hn, wi) > 0.0f) hn = vneg(hn);
                const matp hm = load_mat(s, hi.mat);
                const v3 hwo = vneg(wi);
                suffix = vadd(suffix, vmul(thp, rpt_lights(s, hi.pos, hn, hwo, &hm, bias, &seed)));
                cur_p = hi.pos; cur_n = hn; cur_m = hm; cur_wo = hwo;
            }
            rpt_res cand;
            cand.rc_pos = rc.pos; cand.rc_n = rn; cand.suffix = suffix; cand.prefix = V(1, 1, 1);
            cand.M = 1; cand.W = 0.0f; cand.w_sum = 0.0f; cand.valid = 1.0f;
            float ph = rpt_target(pp, pn, &pm, cp, &cand);
            if (ph <= 0.0f) continue;
            v3 dl = vsub(rc.pos, pp);
            if (vdot(dl, dl) <= 1e-8f) continue;
            float w = ph / fmaxf(src_pdf, 1e-6f);
            w_sum += w;
            Mc++;
            if (w_sum > 0.0f && orc_rand_float(&seed) < w / w_sum) sel = cand;
        }
        int stored = 0;
        if (Mc > 0 && sel.valid > 0.5f) {
            float ph = rpt_target(pp, pn, &pm, cp, &sel);
            if (ph > 1e-10f) {
                sel.M = Mc;
                sel.w_sum = w_sum;
                sel.W = w_sum / ((float)Mc * ph);
                rpt_store(res, npix, i, &sel);
                stored = 1;
            }
        }
        if (!stored) rpt_store_empty(res, npix, i);
    }
}

/* reproject_to_prev_pixel, restir_pt.slang:323-338 */
static int rpt_reproject(const orc_camera* cam, v3 p, uint32_t W, uint32_t H, int* hx, int* hy) {
    const float* M = cam->prev_view_proj;
    float cx = mrow(M, 0, p.x, p.y, p.z, 1.0f), cy = mrow(M, 1, p.x, p.y, p.z, 1.0f),
          cw = mrow(M, 3, p.x, p.y, p.z, 1.0f);
    if (cw <= 1.0e-6f) return 0;
    float nx = cx / cw, ny = cy / cw;
    if (fabsf(nx) > 1.0f || fabsf(ny) > 1.0f) return 0;
    float fx = (nx * 0.5f + 0.5f) * (float)W - 0.5f, fy = (ny * 0.5f + 0.5f) * (float)H - 0.5f;
    *hx = (int)floorf(fx + 0.5f);
    *hy = (int)floorf(fy + 0.5f);
    return *hx >= 0 && *hy >= 0 && *hx < (int)W && *hy < (int)H;
}

/* find_reprojected_history_surface, restir_pt.slang:355-401 (restir_gi.slang:358-433) */
static int rpt_find_history(const orc_camera* cam, const float* hpos_prev, const float* hnrm_prev, uint32_t W,
                            uint32_t H, v3 pp, v3 pn, uint32_t pmi, v3 cp, size_t* hj_out, v3* hp_out) {
    int cx, cy;
    if (!rpt_reproject(cam, pp, W, H, &cx, &cy)) return 0;
    int found = 0;
    float best = 1e30f;
    size_t hj = 0;
    v3 hp = V(0, 0, 0);
    float vdist = fmaxf(vlength(vsub(pp, cp)), 1.0f);
    for (int oy = -1; oy <= 1; ++oy)
        for (int ox = -1; ox <= 1; ++ox) {
            int qx = cx + ox, qy = cy + oy;
            if (qx < 0 || qy < 0 || qx >= (int)W || qy >= (int)H) continue;
            size_t j = (size_t)qy * W + (size_t)qx;
            v3 q, qn; uint32_t qm;
            if (!rpt_surface(hpos_prev, hnrm_prev, j, &q, &qn, &qm)) continue;
            /* is_reprojected_history_compatible, restir_pt.slang:340-353 */
            if (qm != pmi || vdot(pn, qn) < 0.9f) continue;
            float plane = fabsf(vdot(vsub(q, pp), pn));
            float dpos = vlength(vsub(q, pp));
            if (!(plane <= 0.01f * vdist && dpos <= 0.05f * vdist)) continue;
            float err = plane * 4.0f + dpos;
            if (err < best) { best = err; hj = j; hp = q; found = 1; }
        }
    *hj_out = hj;
    *hp_out = hp;
    return found;
}

void orc_rpt_temporal(const orc_scene* s, const orc_camera* cam, const orc_rpt_params* p, const float* gpos,
                      const float* gnrm, float* res, const float* res_prev, const float* hpos_prev,
                      const float* hnrm_prev, int threads) {
    if (!p->temporal_enable) return;
    const uint32_t W = cam->screen_width, H = cam->screen_height;
    const size_t npix = (size_t)W * H;
    const v3 cp = vld(cam->position);
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t ii = 0; ii < (int64_t)npix; ++ii) {
        const size_t i = (size_t)ii;
        const uint32_t px = (uint32_t)(i % W), py = (uint32_t)(i / W);
        v3 pp, pn; uint32_t pmi;
        size_t hj;
        v3 hp;
        if (!rpt_surface(gpos, gnrm, i, &pp, &pn, &pmi) ||
            !rpt_find_history(cam, hpos_prev, hnrm_prev, W, H, pp, pn, pmi, cp, &hj, &hp)) continue;
        rpt_res cur = rpt_load(res, npix, i);
        rpt_res prv = rpt_load(res_prev, npix, hj);
        if (prv.M == 0 || prv.valid < 0.5f) continue;
        uint32_t cap = 20u * p->num_initial_candidates;
        cap = cap < 1u ? 1u : (cap > 8192u ? 8192u : cap);
        prv.M = prv.M < cap ? prv.M : cap;
        float jac = rpt_jacobian(hp, pp, prv.rc_pos, prv.rc_n);
        if (jac <= 0.0f) continue;
        if (!rpt_visible(s, pp, pn, prv.rc_pos, p->ray_bias)) continue;
        const matp pm = load_mat(s, pmi);
        float ph = rpt_target(pp, pn, &pm, cp, &prv);
        float contrib = ph * fminf(prv.W, 50.0f) * (float)prv.M * jac;
        if (contrib <= 0.0f) continue;
        cur.w_sum += contrib;
        uint32_t seed = orc_make_seed(px, py, p->frame_seed + 1u);
        if (cur.w_sum > 0.0f && orc_rand_float(&seed) < contrib / cur.w_sum) {
            cur.rc_pos = prv.rc_pos; cur.rc_n = prv.rc_n; cur.suffix = prv.suffix; cur.prefix = prv.prefix;
            cur.valid = 1.0f;
        }
        uint32_t nM = cur.M + prv.M;
        if (nM > cap) { cur.w_sum *= (float)cap / (float)nM; cur.M = cap; }
        else cur.M = nM;
        if (cur.M > 0 && cur.valid > 0.5f) {
            float py_ = rpt_target(pp, pn, &pm, cp, &cur);
            cur.W = (py_ > 1e-10f) ? (cur.w_sum / ((float)cur.M * py_)) : 0.0f;
            cur.W = fminf(cur.W, 50.0f);
        }
        rpt_store(res, npix, i, &cur);
    }
}

static float fracf_(float x) { return x - floorf(x); }

void orc_rpt_spatial_shade(const orc_scene* s, const orc_camera* cam, const orc_rpt_params* p, uint32_t* accum,
                           float base_cell, float slope, uint32_t cell_count, const float* gpos, const float* gnrm,
                           const float* direct, const float* res, float* out, int threads) {
    const uint32_t W = cam->screen_width, H = cam->screen_height;
    const size_t npix = (size_t)W * H;
    const v3 cp = vld(cam->position);
    float cpa[3]; vst(cpa, cp);
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t ii = 0; ii < (int64_t)npix; ++ii) {
        const size_t i = (size_t)ii;
        const uint32_t px = (uint32_t)(i % W), py = (uint32_t)(i / W);
        v3 L = V(0, 0, 0);
        if (p->include_direct_lighting != 0 && direct) L = vld(direct + i * 4);
        v3 pp, pn; uint32_t pmi;
        if (!rpt_surface(gpos, gnrm, i, &pp, &pn, &pmi)) {
            if (p->include_direct_lighting == 0) L = vadd(L, sky(center_ray_dir(cam, px, py)));
        } else {
            const matp pm = load_mat(s, pmi);
            if (pm.emissive.x > 0.0f || pm.emissive.y > 0.0f || pm.emissive.z > 0.0f) {
                L = vadd(L, pm.emissive);
            } else {
                rpt_res r = rpt_load(res, npix, i);
                if (p->spatial_enable && p->spatial_num_neighbors > 0 && r.valid > 0.5f) {
                    float base_phase = fracf_((float)orc_pcg_hash(px ^ (py * 65537u)) * 2.3283064365386963e-10f);
                    float dist_scale = fmaxf(vlength(vsub(pp, cp)), 1.0f);
                    for (uint32_t k = 0; k < p->spatial_num_neighbors; ++k) {
                        float u = ((float)k + 0.5f) / (float)p->spatial_num_neighbors;
                        float angle = fracf_(base_phase + 0.61803398875f * (float)k) * K_TWO_PI;
                        float radius = sqrtf(u) * (float)p->spatial_radius;
                        float sa, ca; orc_sincosf(angle, &sa, &ca);
                        int nx = (int)px + (int)(radius * ca), ny = (int)py + (int)(radius * sa);
                        nx = nx < 0 ? 0 : (nx > (int)W - 1 ? (int)W - 1 : nx);
                        ny = ny < 0 ? 0 : (ny > (int)H - 1 ? (int)H - 1 : ny);
                        if (nx == (int)px && ny == (int)py) continue;
                        size_t j = (size_t)ny * W + (size_t)nx;
                        /* is_spatial_neighbor_compatible, restir_pt.slang:403-421 */
                        const float* q4 = gpos + j * 4;
                        if (q4[3] < 0.0f) continue;
                        v3 q = vld(q4), qn = vnorm(vld(gnrm + j * 4));
                        if (vdot(qn, pn) < 0.85f) continue;
                        if (!(fabsf(vdot(vsub(q, pp), pn)) <= 0.02f * dist_scale)) continue;
                        rpt_res nb = rpt_load(res, npix, j);
                        if (nb.M == 0 || nb.valid < 0.5f) continue;
                        float jac = rpt_jacobian(q, pp, nb.rc_pos, nb.rc_n);
                        if (jac <= 0.0f) continue;
                        float contrib = rpt_target(pp, pn, &pm, cp, &nb) * nb.W * (float)nb.M * jac;
                        if (contrib <= 0.0f) continue;
                        r.w_sum += contrib;
                        r.M += nb.M;
                        float accept = fracf_(base_phase + 0.754877666f * (float)(k + 1u));
                        if (r.w_sum > 0.0f && accept < contrib / r.w_sum) {
                            r.rc_pos = nb.rc_pos; r.rc_n = nb.rc_n; r.suffix = nb.suffix; r.prefix = nb.prefix;
                        }
                    }
                    if (r.M > 0) {
                        float ph = rpt_target(pp, pn, &pm, cp, &r);
                        r.W = (ph > 1e-10f) ? (r.w_sum / ((float)r.M * ph)) : 0.0f;
                    }
                }
                if (r.M > 0 && r.W > 0.0f && r.valid > 0.5f) {
                    v3 ind = vscale(rpt_contribution(pp, pn, &pm, cp, &r), r.W);
                    if (lum709(ind) > 0.0f && rpt_visible(s, pp, pn, r.rc_pos, p->ray_bias))
                        L = vadd(L, vscale(ind, p->indirect_strength));
                    if (p->guiding_enable && accum) {
                        v3 to = vsub(r.rc_pos, pp);
                        float d2 = vdot(to, to);
                        if (d2 > 1e-8f) {
                            float ppa[3], dir[3];
                            vst(ppa, pp); vst(dir, vscale(to, 1.0f / sqrtf(d2)));
                            uint32_t cell = orc_guiding_cell_index(ppa, pmi, cpa, base_cell, slope, cell_count);
                            uint32_t* a = accum + (size_t)cell * 4;
                            uint32_t dx = (uint32_t)(int32_t)roundf(dir[0] * 1048576.0f);
                            uint32_t dy = (uint32_t)(int32_t)roundf(dir[1] * 1048576.0f);
                            uint32_t dz = (uint32_t)(int32_t)roundf(dir[2] * 1048576.0f);
#pragma omp atomic
                            a[0] += dx;
#pragma omp atomic
                            a[1] += dy;
#pragma omp atomic
                            a[2] += dz;
#pragma omp atomic
                            a[3] += 1u;
                        }
                    }
                }
            }
        }
        float* o = out + i * 4;
        o[0] = L.x; o[1] = L.y; o[2] = L.z; o[3] = 1.0f;
    }
}

/* ========================================================================== */
/* ReSTIR GI, one-bounce reconnection reuse -- restir_gi.slang:143-1013.  The
 * helpers are ReSTIR PT's (the shaders define them identically); here the target is
 * brdf * radiance * cos, the source pdf cos / pi toward the hit, the kept normal the
 * interpolated one, and M = 0 marks an empty reservoir (three planes).            */
/* ========================================================================== */
typedef struct { v3 pos, rad, n; float w_sum, W; uint32_t M; } rgi_res;

static rgi_res rgi_load(const float* res, size_t npix, size_t i) {
    const float* p = res + i * 4; const float* r = res + (npix + i) * 4; const float* m = res + (2 * npix + i) * 4;
    rgi_res s;
    s.pos = vld(p); s.w_sum = p[3];
    s.rad = vld(r); s.W = r[3];
    s.n = vld(m); s.M = asuint(m[3]);
    return s;
}
static void rgi_store(float* res, size_t npix, size_t i, const rgi_res* s) {
    float* p = res + i * 4; float* r = res + (npix + i) * 4; float* m = res + (2 * npix + i) * 4;
    vst(p, s->pos); p[3] = s->w_sum;
    vst(r, s->rad); r[3] = s->W;
    vst(m, s->n); m[3] = asfloat(s->M);
}
static void rgi_store_empty(float* res, size_t npix, size_t i) {
    for (int k = 0; k < 3; ++k) memset(res + (k * npix + i) * 4, 0, 16);
}
/* evaluate_indirect_radiance / compute_target_function, restir_gi.slang:487-537 */
static v3 rgi_indirect(v3 pp, v3 pn, const matp* pm, v3 sp, v3 srad, v3 cam) {
    v3 to = vsub(sp, pp);
    float d2 = vdot(to, to);
    if (d2 <= 1e-8f) return V(0, 0, 0);
    v3 wi = vscale(to, 1.0f / sqrtf(d2));
    float cx = fmaxf(vdot(pn, wi), 0.0f);
    if (cx <= 0.0f) return V(0, 0, 0);
    v3 wo = vnorm(vsub(cam, pp));
    return vscale(vmul(mat_eval(pm, pn, wo, wi), srad), cx);
}
static float rgi_target(v3 pp, v3 pn, const matp* pm, v3 sp, v3 srad, v3 cam) {
    return lum709(rgi_indirect(pp, pn, pm, sp, srad, cam));
}
/* sample_primary_pdf, restir_gi.slang:510-522 */
static float rgi_primary_pdf(v3 pp, v3 pn, v3 sp) {
    v3 d = vsub(sp, pp);
    float d2 = vdot(d, d);
    if (d2 <= 1e-8f) return 0.0f;
    v3 wi = vscale(d, 1.0f / sqrtf(d2));
    float cx = fmaxf(vdot(pn, wi), 0.0f);
    return fmaxf(cx * K_INV_PI, 0.0f);
}

void orc_rgi_initial(const orc_scene* s, const orc_camera* cam, const orc_rgi_params* p, const float* gpos,
                     const float* gnrm, float* res, float* hpos, float* hnrm, int threads) {
    const uint32_t W = cam->screen_width, H = cam->screen_height;
    const size_t npix = (size_t)W * H;
    const v3 cp = vld(cam->position);
    const int hit_emission = s->ntl == 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t ii = 0; ii < (int64_t)npix; ++ii) {
        const size_t i = (size_t)ii;
        const uint32_t px = (uint32_t)(i % W), py = (uint32_t)(i / W);
        v3 pp, pn; uint32_t pmi;
        if (!rpt_surface(gpos, gnrm, i, &pp, &pn, &pmi)) {
            float* h = hpos + i * 4; h[0] = h[1] = h[2] = 0.0f; h[3] = -1.0f;
            memset(hnrm + i * 4, 0, 16);
            rgi_store_empty(res, npix, i);
            continue;
        }
        memcpy(hpos + i * 4, gpos + i * 4, 16);
        memcpy(hnrm + i * 4, gnrm + i * 4, 16);
        const matp pm = load_mat(s, pmi);
        const float bias = p->ray_bias;
        uint32_t seed = orc_make_seed(px, py, p->frame_seed);
        rgi_res sel; memset(&sel, 0, sizeof(sel));
        float w_sum = 0.0f;
        uint32_t Mc = 0;
        for (uint32_t c = 0; c < p->num_initial_candidates; ++c) {
            float u0 = orc_rand_float(&seed), u1 = orc_rand_float(&seed), l[3], pdf_dir;
            orc_sample_cosine_hemisphere(u0, u1, l, &pdf_dir);
            v3 T, B; build_onb(pn, &T, &B);
            v3 dir = to_world(vld(l), T, B, pn);
            if (vdot(pn, dir) <= 0.0f) continue;
            orc_hit h = trace_bvh_cull(s, vadd(pp, vscale(pn, bias)), dir, bias, 1e9f, g_restir_cull);
            if (h.prim == ~0u) continue;
            hitinfo hi = fetch_hit(s, h);
            /* evaluate_sample_radiance, restir_gi.slang:539-551 */
            v3 wo = vneg(dir);
            v3 n = hi.normal;
            if (vdot(n, wo) < 0.0f) n = vneg(n);
            const matp m = load_mat(s, hi.mat);
            v3 srad = hit_emission ? m.emissive : V(0, 0, 0);
            srad = vadd(srad, rpt_lights(s, hi.pos, n, wo, &m, bias, &seed));
            float tgt = rgi_target(pp, pn, &pm, hi.pos, srad, cp);
            if (tgt <= 0.0f) continue;
            float pdf = rgi_primary_pdf(pp, pn, hi.pos);
            if (pdf <= 1e-6f) continue;
            float w = tgt / pdf;
            w_sum += w;
            Mc++;
            if (w_sum > 0.0f && orc_rand_float(&seed) < w / w_sum) {
                sel.pos = hi.pos; sel.rad = srad; sel.n = hi.normal;
            }
        }
        int stored = 0;
        if (Mc > 0) {
            float ph = rgi_target(pp, pn, &pm, sel.pos, sel.rad, cp);
            if (ph > 1e-10f) {
                sel.w_sum = w_sum;
                sel.W = w_sum / ((float)Mc * ph);
                sel.M = Mc;
                rgi_store(res, npix, i, &sel);
                stored = 1;
            }
        }
        if (!stored) rgi_store_empty(res, npix, i);
    }
}

void orc_rgi_temporal(const orc_scene* s, const orc_camera* cam, const orc_rgi_params* p, const float* gpos,
                      const float* gnrm, float* res, const float* res_prev, const float* hpos_prev,
                      const float* hnrm_prev, int threads) {
    if (!p->temporal_enable) return;
    const uint32_t W = cam->screen_width, H = cam->screen_height;
    const size_t npix = (size_t)W * H;
    const v3 cp = vld(cam->position);
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t ii = 0; ii < (int64_t)npix; ++ii) {
        const size_t i = (size_t)ii;
        const uint32_t px = (uint32_t)(i % W), py = (uint32_t)(i / W);
        v3 pp, pn; uint32_t pmi;
        size_t hj;
        v3 hp;
        if (!rpt_surface(gpos, gnrm, i, &pp, &pn, &pmi) ||
            !rpt_find_history(cam, hpos_prev, hnrm_prev, W, H, pp, pn, pmi, cp, &hj, &hp)) continue;
        rgi_res cur = rgi_load(res, npix, i);
        rgi_res prv = rgi_load(res_prev, npix, hj);
        if (prv.M == 0) continue;
        uint32_t cap = 20u * p->num_initial_candidates;
        cap = cap < 1u ? 1u : (cap > 8192u ? 8192u : cap);
        prv.M = prv.M < cap ? prv.M : cap;
        float jac = rpt_jacobian(hp, pp, prv.pos, prv.n);
        if (jac <= 0.0f) continue;
        if (!rpt_visible(s, pp, pn, prv.pos, p->ray_bias)) continue;
        const matp pm = load_mat(s, pmi);
        float ph = rgi_target(pp, pn, &pm, prv.pos, prv.rad, cp);
        float contrib = ph * fminf(prv.W, 50.0f) * (float)prv.M * jac;
        if (contrib <= 0.0f) continue;
        cur.w_sum += contrib;
        uint32_t seed = orc_make_seed(px, py, p->frame_seed + 1u);
        if (cur.w_sum > 0.0f && orc_rand_float(&seed) < contrib / cur.w_sum) {
            cur.pos = prv.pos; cur.rad = prv.rad; cur.n = prv.n;
        }
        uint32_t nM = cur.M + prv.M;
        if (nM > cap) { cur.w_sum *= (float)cap / (float)nM; cur.M = cap; }
        else cur.M = nM;
        float Wn = 0.0f;
        if (cur.M > 0) {
            float py_ = rgi_target(pp, pn, &pm, cur.pos, cur.rad, cp);
            Wn = (py_ > 1e-10f) ? (cur.w_sum / ((float)cur.M * py_)) : 0.0f;
            Wn = fminf(Wn, 50.0f);
        }
        cur.W = Wn;
        rgi_store(res, npix, i, &cur);
    }
}

void orc_rgi_spatial_shade(const orc_scene* s, const orc_camera* cam, const orc_rgi_params* p, const float* gpos,
                           const float* gnrm, const float* direct, const float* res, float* out, int threads) {
    const uint32_t W = cam->screen_width, H = cam->screen_height;
    const size_t npix = (size_t)W * H;
    const v3 cp = vld(cam->position);
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t ii = 0; ii < (int64_t)npix; ++ii) {
        const size_t i = (size_t)ii;
        const uint32_t px = (uint32_t)(i % W), py = (uint32_t)(i / W);
        v3 L = V(0, 0, 0);
        if (p->include_direct_lighting != 0 && direct) L = vld(direct + i * 4);
        v3 pp, pn; uint32_t pmi;
        if (!rpt_surface(gpos, gnrm, i, &pp, &pn, &pmi)) {
            if (p->include_direct_lighting == 0) L = vadd(L, sky(center_ray_dir(cam, px, py)));
        } else {
            const matp pm = load_mat(s, pmi);
            if (pm.emissive.x > 0.0f || pm.emissive.y > 0.0f || pm.emissive.z > 0.0f) {
                L = vadd(L, pm.emissive);
            } else {
                rgi_res r = rgi_load(res, npix, i);
                if (p->spatial_enable && p->spatial_num_neighbors > 0) {
                    float base_phase = fracf_((float)orc_pcg_hash(px ^ (py * 65537u)) * 2.3283064365386963e-10f);
                    float dist_scale = fmaxf(vlength(vsub(pp, cp)), 1.0f);
                    for (uint32_t k = 0; k < p->spatial_num_neighbors; ++k) {
                        float u = ((float)k + 0.5f) / (float)p->spatial_num_neighbors;
                        float angle = fracf_(base_phase + 0.61803398875f * (float)k) * K_TWO_PI;
                        float radius = sqrtf(u) * (float)p->spatial_radius;
                        float sa, ca; orc_sincosf(angle, &sa, &ca);
                        int nx = (int)px + (int)(radius * ca), ny = (int)py + (int)(radius * sa);
                        nx = nx < 0 ? 0 : (nx > (int)W - 1 ? (int)W - 1 : nx);
                        ny = ny < 0 ? 0 : (ny > (int)H - 1 ? (int)H - 1 : ny);
                        if (nx == (int)px && ny == (int)py) continue;
                        size_t j = (size_t)ny * W + (size_t)nx;
                        /* is_spatial_neighbor_compatible, restir_gi.slang:435-459 */
                        const float* q4 = gpos + j * 4;
                        if (q4[3] < 0.0f) continue;
                        v3 q = vld(q4), qn = vnorm(vld(gnrm + j * 4));
                        if (vdot(qn, pn) < 0.85f) continue;
                        if (!(fabsf(vdot(vsub(q, pp), pn)) <= 0.02f * dist_scale)) continue;
                        rgi_res nb = rgi_load(res, npix, j);
                        if (nb.M == 0) continue;
                        float jac = rpt_jacobian(q, pp, nb.pos, nb.n);
                        if (jac <= 0.0f) continue;
                        float contrib = rgi_target(pp, pn, &pm, nb.pos, nb.rad, cp) * nb.W * (float)nb.M * jac;
                        if (contrib <= 0.0f) continue;
                        r.w_sum += contrib;
                        r.M += nb.M;
                        float accept = fracf_(base_phase + 0.754877666f * (float)(k + 1u));
                        if (r.w_sum > 0.0f && accept < contrib / r.w_sum) { r.pos = nb.pos; r.rad = nb.rad; r.n = nb.n; }
                    }
                    if (r.M > 0) {
                        float ph = rgi_target(pp, pn, &pm, r.pos, r.rad, cp);
                        r.W = (ph > 1e-10f) ? (r.w_sum / ((float)r.M * ph)) : 0.0f;
                    }
                }
                if (r.M > 0 && r.W > 0.0f) {
                    v3 ind = vscale(rgi_indirect(pp, pn, &pm, r.pos, r.rad, cp), r.W);
                    if (lum709(ind) > 0.0f && rpt_visible(s, pp, pn, r.pos, p->ray_bias))
                        L = vadd(L, vscale(ind, p->indirect_strength));
                }
            }
        }
        float* o = out + i * 4;
        o[0] = L.x; o[1] = L.y; o[2] = L.z; o[3] = 1.0f;
    }
}

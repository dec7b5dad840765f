// rrh_scenegen.cpp -- seeded scene generators.
//
// RRH_GEN_CORNELL_REF restates Scene::build_cornell_box (src/scene/Scene.cpp:144-251)
// with add_quad / add_box (:94-141): same corners, winding, normals, materials,
// instance transforms (translate * rotate about +y) and the point light.  The other
// generators are new (the reference has no procedural scenes); all draw from
// pcg_hash so a seed reproduces the scene on any host.
#include <cmath>
#include <cstring>
#include <utility>

#include "rrh_internal.h"

namespace rrh {
namespace {

struct Rng {
    uint32_t s;
    explicit Rng(uint32_t seed) : s(seed * 2654435761u + 0x9e3779b9u) {}
    uint32_t next() {
        uint32_t state = s * 747796405u + 2891336453u;
        uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
        s = (word >> 22u) ^ word;
        return s;
    }
    float uf() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
    float range(float a, float b) { return a + (b - a) * uf(); }
    float logu(float a, float b) { return std::exp(std::log(a) + (std::log(b) - std::log(a)) * uf()); }
};

struct V3 { float x, y, z; };
V3 v3(float x, float y, float z) { return V3{x, y, z}; }

rrh_vertex make_v(V3 p, V3 n, float u, float v) {
    rrh_vertex r;
    r.pos[0] = p.x; r.pos[1] = p.y; r.pos[2] = p.z;
    r.normal[0] = n.x; r.normal[1] = n.y; r.normal[2] = n.z;
    r.uv[0] = u; r.uv[1] = v;
    r.tangent[0] = 1; r.tangent[1] = 0; r.tangent[2] = 0; r.tangent[3] = 1;
    return r;
}

// Scene.cpp:94-115: four corners, triangles (0,1,2) and (0,2,3)
void add_quad(std::vector<rrh_vertex>& vs, std::vector<uint32_t>& is, V3 a, V3 b, V3 c, V3 d, V3 n) {
    uint32_t base = (uint32_t)vs.size();
    vs.push_back(make_v(a, n, 0, 0));
    vs.push_back(make_v(b, n, 1, 0));
    vs.push_back(make_v(c, n, 1, 1));
    vs.push_back(make_v(d, n, 0, 1));
    uint32_t t[6] = {base, base + 1, base + 2, base, base + 2, base + 3};
    is.insert(is.end(), t, t + 6);
}

// a quad whose two triangles face its normal n: geometric normal (v1 - v0) x (v2 - v0) along
// n, the front face of back-face culling (the G-buffer and ReSTIR PT cull back faces).  The
// reference's add_quad keeps the corner order it is given, and several quads of its Cornell box
// face away from their normals; the reference scene keeps that, the new generators face out.
void add_quad_facing(std::vector<rrh_vertex>& vs, std::vector<uint32_t>& is, V3 a, V3 b, V3 c, V3 d, V3 n) {
    const float ux = b.x - a.x, uy = b.y - a.y, uz = b.z - a.z, vx = c.x - a.x, vy = c.y - a.y, vz = c.z - a.z;
    const float gx = uy * vz - uz * vy, gy = uz * vx - ux * vz, gz = ux * vy - uy * vx;
    if (gx * n.x + gy * n.y + gz * n.z < 0.0f) add_quad(vs, is, a, d, c, b, n);
    else add_quad(vs, is, a, b, c, d, n);
}

// a box whose faces all face outward
void add_box_facing(std::vector<rrh_vertex>& vs, std::vector<uint32_t>& is, V3 mn, V3 mx) {
    V3 a{mn.x, mn.y, mn.z}, b{mx.x, mn.y, mn.z}, c{mx.x, mn.y, mx.z}, d{mn.x, mn.y, mx.z};
    V3 e{mn.x, mx.y, mn.z}, f{mx.x, mx.y, mn.z}, g{mx.x, mx.y, mx.z}, h{mn.x, mx.y, mx.z};
    add_quad_facing(vs, is, d, c, b, a, v3(0, -1, 0));
    add_quad_facing(vs, is, e, f, g, h, v3(0, 1, 0));
    add_quad_facing(vs, is, d, h, g, c, v3(0, 0, 1));
    add_quad_facing(vs, is, b, f, e, a, v3(0, 0, -1));
    add_quad_facing(vs, is, a, e, h, d, v3(-1, 0, 0));
    add_quad_facing(vs, is, c, g, f, b, v3(1, 0, 0));
}

// Scene.cpp:118-141
void add_box(std::vector<rrh_vertex>& vs, std::vector<uint32_t>& is, V3 mn, V3 mx) {
    V3 a{mn.x, mn.y, mn.z}, b{mx.x, mn.y, mn.z}, c{mx.x, mn.y, mx.z}, d{mn.x, mn.y, mx.z};
    V3 e{mn.x, mx.y, mn.z}, f{mx.x, mx.y, mn.z}, g{mx.x, mx.y, mx.z}, h{mn.x, mx.y, mx.z};
    add_quad(vs, is, d, c, b, a, v3(0, -1, 0));
    add_quad(vs, is, e, f, g, h, v3(0, 1, 0));
    add_quad(vs, is, d, h, g, c, v3(0, 0, 1));
    add_quad(vs, is, b, f, e, a, v3(0, 0, -1));
    add_quad(vs, is, a, e, h, d, v3(-1, 0, 0));
    add_quad(vs, is, c, g, f, b, v3(1, 0, 0));
}

rrh_material material(float r, float g, float b, float metallic, float roughness, float er = 0, float eg = 0,
                      float eb = 0, uint32_t flags = 0) {
    rrh_material m;
    std::memset(&m, 0, sizeof(m));
    m.base_color[0] = r; m.base_color[1] = g; m.base_color[2] = b; m.base_color[3] = 1.0f;
    m.metallic = metallic; m.roughness = roughness;
    m.emissive[0] = er; m.emissive[1] = eg; m.emissive[2] = eb;
    m.ior = 1.5f;
    m.albedo_tex_idx = m.normal_tex_idx = m.metallic_roughness_tex_idx = m.emissive_tex_idx = 0xffffffffu;
    m.flags = flags;
    return m;
}

void identity(float* T) {
    std::memset(T, 0, 64);
    T[0] = T[5] = T[10] = T[15] = 1.0f;
}

// append one mesh (vertex_base + index_byte_offset bookkeeping of Scene::upload, Scene.cpp:283-318)
uint32_t add_mesh(SceneData& s, const std::vector<rrh_vertex>& vs, const std::vector<uint32_t>& is, uint32_t mat,
                  const float* T = nullptr) {
    rrh_mesh m;
    std::memset(&m, 0, sizeof(m));
    m.vertex_base = (uint32_t)s.verts.size();
    m.vertex_count = (uint32_t)vs.size();
    m.index_byte_offset = (uint32_t)(s.idx.size() * 4);
    m.index_count = (uint32_t)is.size();
    m.material_index = mat;
    s.verts.insert(s.verts.end(), vs.begin(), vs.end());
    s.idx.insert(s.idx.end(), is.begin(), is.end());
    s.meshes.push_back(m);
    rrh_instance in;
    std::memset(&in, 0, sizeof(in));
    if (T) std::memcpy(in.transform, T, 64);
    else identity(in.transform);
    in.mesh_index = (uint32_t)s.meshes.size() - 1;
    s.insts.push_back(in);
    return in.mesh_index;
}

// glm::translate(I, t) * glm::rotate(I, radians(deg), +y), stored row-major
void trs_y(float* T, float tx, float ty, float tz, float deg) {
    float a = deg * 0.01745329251994329576923690768489f;
    float c = std::cos(a), s = std::sin(a);
    identity(T);
    T[0] = c; T[2] = s; T[8] = -s; T[10] = c;
    T[3] = tx; T[7] = ty; T[11] = tz;
}

void cornell_ref(SceneData& s) {
    uint32_t white = (uint32_t)s.mats.size(); s.mats.push_back(material(0.73f, 0.73f, 0.73f, 0.0f, 1.0f));
    uint32_t red = (uint32_t)s.mats.size(); s.mats.push_back(material(0.65f, 0.05f, 0.05f, 0.0f, 1.0f));
    uint32_t green = (uint32_t)s.mats.size(); s.mats.push_back(material(0.12f, 0.45f, 0.15f, 0.0f, 1.0f));
    uint32_t light = (uint32_t)s.mats.size(); s.mats.push_back(material(1, 1, 1, 0.0f, 1.0f, 15, 15, 15));
    std::vector<rrh_vertex> v; std::vector<uint32_t> i;
    auto quad_mesh = [&](V3 a, V3 b, V3 c, V3 d, V3 n, uint32_t m) {
        v.clear(); i.clear(); add_quad(v, i, a, b, c, d, n); add_mesh(s, v, i, m);
    };
    quad_mesh(v3(-1, -1, -1), v3(1, -1, -1), v3(1, 1, -1), v3(-1, 1, -1), v3(0, 0, 1), white);   // back
    quad_mesh(v3(-1, -1, 1), v3(1, -1, 1), v3(1, -1, -1), v3(-1, -1, -1), v3(0, 1, 0), white);   // floor
    quad_mesh(v3(-1, 1, -1), v3(1, 1, -1), v3(1, 1, 1), v3(-1, 1, 1), v3(0, -1, 0), white);      // ceiling
    quad_mesh(v3(-1, -1, -1), v3(-1, -1, 1), v3(-1, 1, 1), v3(-1, 1, -1), v3(1, 0, 0), red);     // left
    quad_mesh(v3(1, -1, 1), v3(1, -1, -1), v3(1, 1, -1), v3(1, 1, 1), v3(-1, 0, 0), green);      // right
    quad_mesh(v3(-0.25f, 0.999f, -0.25f), v3(0.25f, 0.999f, -0.25f), v3(0.25f, 0.999f, 0.25f),
              v3(-0.25f, 0.999f, 0.25f), v3(0, -1, 0), light);                                  // light quad
    float T[16];
    v.clear(); i.clear(); add_box(v, i, v3(-0.25f, -1.0f, -0.25f), v3(0.25f, -0.5f, 0.25f));
    trs_y(T, 0.35f, 0.0f, -0.35f, 18.0f);
    add_mesh(s, v, i, white, T);
    v.clear(); i.clear(); add_box(v, i, v3(-0.25f, -1.0f, -0.25f), v3(0.25f, 0.0f, 0.25f));
    trs_y(T, -0.35f, 0.0f, -0.2f, -15.0f);
    add_mesh(s, v, i, white, T);
    rrh_light l;
    std::memset(&l, 0, sizeof(l));
    l.type = 0;
    l.position[0] = 0.0f; l.position[1] = 0.9f; l.position[2] = 0.0f;
    l.emission[0] = 15.0f; l.emission[1] = 15.0f; l.emission[2] = 12.0f;
    l.intensity = 1.0f;
    s.lights.push_back(l);
    s.flags = 0;
    s.eye[0] = 0; s.eye[1] = 0; s.eye[2] = 3;
    s.center[0] = 0; s.center[1] = 0; s.center[2] = 0;
    s.fov = 45.0f;
}

// 36 triangles: 5 walls + light quad + 2 random axis-aligned blocks; Lambert albedos U[.1,.9]
void cornell_random(SceneData& s, uint32_t seed) {
    Rng r(seed);
    auto lam = [&]() {
        float a = r.range(0.1f, 0.9f), b = r.range(0.1f, 0.9f), c = r.range(0.1f, 0.9f);
        s.mats.push_back(material(a, b, c, 0.0f, 1.0f, 0, 0, 0, RRH_MAT_PURE_LAMBERT));
        return (uint32_t)s.mats.size() - 1;
    };
    std::vector<rrh_vertex> v; std::vector<uint32_t> i;
    auto quad_mesh = [&](V3 a, V3 b, V3 c, V3 d, V3 n, uint32_t m) {
        v.clear(); i.clear(); add_quad_facing(v, i, a, b, c, d, n); add_mesh(s, v, i, m);
    };
    quad_mesh(v3(-1, -1, -1), v3(1, -1, -1), v3(1, 1, -1), v3(-1, 1, -1), v3(0, 0, 1), lam());
    quad_mesh(v3(-1, -1, 1), v3(1, -1, 1), v3(1, -1, -1), v3(-1, -1, -1), v3(0, 1, 0), lam());
    quad_mesh(v3(-1, 1, -1), v3(1, 1, -1), v3(1, 1, 1), v3(-1, 1, 1), v3(0, -1, 0), lam());
    quad_mesh(v3(-1, -1, -1), v3(-1, -1, 1), v3(-1, 1, 1), v3(-1, 1, -1), v3(1, 0, 0), lam());
    quad_mesh(v3(1, -1, 1), v3(1, -1, -1), v3(1, 1, -1), v3(1, 1, 1), v3(-1, 0, 0), lam());
    s.mats.push_back(material(1, 1, 1, 0.0f, 1.0f, 15, 15, 15, RRH_MAT_PURE_LAMBERT));
    quad_mesh(v3(-0.25f, 0.999f, -0.25f), v3(0.25f, 0.999f, -0.25f), v3(0.25f, 0.999f, 0.25f),
              v3(-0.25f, 0.999f, 0.25f), v3(0, -1, 0), (uint32_t)s.mats.size() - 1);
    for (int b = 0; b < 2; ++b) {
        float sx = r.range(0.25f, 0.6f), sz = r.range(0.25f, 0.6f), h = r.range(0.3f, 1.2f);
        float cx = (b == 0) ? r.range(-0.7f, -0.05f) + 0.0f : r.range(0.05f, 0.7f);
        float cz = r.range(-0.6f, 0.4f);
        float x0 = std::fmax(-0.98f, cx - 0.5f * sx), x1 = std::fmin(0.98f, cx + 0.5f * sx);
        float z0 = std::fmax(-0.98f, cz - 0.5f * sz), z1 = std::fmin(0.98f, cz + 0.5f * sz);
        v.clear(); i.clear();
        add_box_facing(v, i, v3(x0, -1.0f, z0), v3(x1, -1.0f + h, z1));
        add_mesh(s, v, i, lam());
    }
    s.flags = RRH_SCENE_TRI_LIGHTS;
    s.eye[0] = 0; s.eye[1] = 0; s.eye[2] = 3;
    s.center[0] = 0; s.center[1] = 0; s.center[2] = 0;
    s.fov = 45.0f;
}

// ---- value-noise fBm ------------------------------------------------------------
float hash01(int x, int z, uint32_t seed, int oct) {
    uint32_t h = (uint32_t)x * 73856093u ^ (uint32_t)z * 19349663u ^ seed * 83492791u ^ (uint32_t)oct * 2654435761u;
    h = h * 747796405u + 2891336453u;
    h = ((h >> ((h >> 28u) + 4u)) ^ h) * 277803737u;
    h = (h >> 22u) ^ h;
    return (float)(h >> 8) * (1.0f / 16777216.0f);
}
float vnoise(float x, float z, uint32_t seed, int oct) {
    int ix = (int)std::floor(x), iz = (int)std::floor(z);
    float fx = x - (float)ix, fz = z - (float)iz;
    float ux = fx * fx * (3.0f - 2.0f * fx), uz = fz * fz * (3.0f - 2.0f * fz);
    float a = hash01(ix, iz, seed, oct), b = hash01(ix + 1, iz, seed, oct);
    float c = hash01(ix, iz + 1, seed, oct), d = hash01(ix + 1, iz + 1, seed, oct);
    return (a + (b - a) * ux) + ((c + (d - c) * ux) - (a + (b - a) * ux)) * uz;
}
float fbm(float x, float z, uint32_t seed) {
    float amp = 1.0f, f = 1.0f, sum = 0.0f;
    for (int o = 0; o < 6; ++o) {
        sum += amp * (vnoise(x * f, z * f, seed, o) - 0.5f);
        amp *= 0.5f;
        f *= 2.0f;
    }
    return sum;
}

// icosphere with `sub` subdivisions (20 * 4^sub triangles)
void icosphere(int sub, std::vector<V3>& P, std::vector<uint32_t>& I) {
    const float t = (1.0f + std::sqrt(5.0f)) * 0.5f;
    P = {v3(-1, t, 0), v3(1, t, 0), v3(-1, -t, 0), v3(1, -t, 0), v3(0, -1, t), v3(0, 1, t),
         v3(0, -1, -t), v3(0, 1, -t), v3(t, 0, -1), v3(t, 0, 1), v3(-t, 0, -1), v3(-t, 0, 1)};
    I = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
         3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    auto norm = [](V3 a) { float l = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); return v3(a.x / l, a.y / l, a.z / l); };
    for (V3& p : P) p = norm(p);
    for (int s = 0; s < sub; ++s) {
        std::vector<uint32_t> J;
        std::vector<std::pair<uint64_t, uint32_t>> cache;
        auto mid = [&](uint32_t a, uint32_t b) {
            uint64_t key = a < b ? ((uint64_t)a << 32 | b) : ((uint64_t)b << 32 | a);
            for (auto& kv : cache)
                if (kv.first == key) return kv.second;
            V3 m = norm(v3(0.5f * (P[a].x + P[b].x), 0.5f * (P[a].y + P[b].y), 0.5f * (P[a].z + P[b].z)));
            P.push_back(m);
            cache.push_back({key, (uint32_t)P.size() - 1});
            return (uint32_t)P.size() - 1;
        };
        for (size_t k = 0; k < I.size(); k += 3) {
            uint32_t a = I[k], b = I[k + 1], c = I[k + 2];
            uint32_t ab = mid(a, b), bc = mid(b, c), ca = mid(c, a);
            uint32_t tri[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            J.insert(J.end(), tri, tri + 12);
        }
        I.swap(J);
    }
}

// 512^2-cell fBm heightfield + 2000 icospheres (subdivision 2) + 4096 emissive triangles
void heightfield(SceneData& s, uint32_t seed, float scale) {
    Rng r(seed);
    int N = std::max(8, (int)std::lround(512.0f * scale));
    int nsph = std::max(4, (int)std::lround(2000.0f * scale));
    int nem = std::max(4, (int)std::lround(4096.0f * scale));
    const float ext = 100.0f, amp = 12.0f;
    auto height = [&](float x, float z) { return amp * fbm(x * 0.03f, z * 0.03f, seed); };
    // terrain
    uint32_t mat_ground = (uint32_t)s.mats.size();
    s.mats.push_back(material(0.45f, 0.42f, 0.36f, 0.0f, 0.9f));
    {
        std::vector<rrh_vertex> v; std::vector<uint32_t> i;
        v.reserve((size_t)(N + 1) * (N + 1));
        float h = ext / (float)N;
        for (int z = 0; z <= N; ++z)
            for (int x = 0; x <= N; ++x) {
                float px = -0.5f * ext + x * h, pz = -0.5f * ext + z * h;
                float py = height(px, pz);
                float dx = height(px + 0.05f, pz) - height(px - 0.05f, pz);
                float dz = height(px, pz + 0.05f) - height(px, pz - 0.05f);
                V3 n = v3(-dx, 0.2f, -dz);
                float l = std::sqrt(n.x * n.x + n.y * n.y + n.z * n.z);
                v.push_back(make_v(v3(px, py, pz), v3(n.x / l, n.y / l, n.z / l), (float)x / N, (float)z / N));
            }
        for (int z = 0; z < N; ++z)
            for (int x = 0; x < N; ++x) {
                uint32_t a = z * (N + 1) + x, b = a + 1, c = a + (N + 1), d = c + 1;
                uint32_t t[6] = {a, c, b, b, c, d};
                i.insert(i.end(), t, t + 6);
            }
        add_mesh(s, v, i, mat_ground);
    }
    // sphere materials
    const int nmat = 16;
    uint32_t mat0 = (uint32_t)s.mats.size();
    for (int k = 0; k < nmat; ++k)
        s.mats.push_back(material(r.range(0.1f, 0.9f), r.range(0.1f, 0.9f), r.range(0.1f, 0.9f),
                                  r.uf() < 0.25f ? 1.0f : 0.0f, r.range(0.3f, 1.0f)));
    std::vector<V3> P; std::vector<uint32_t> I;
    icosphere(2, P, I);
    const int tps = (int)I.size() / 3;
    // pick the emissive (sphere, triangle) pairs
    std::vector<uint8_t> em((size_t)nsph * tps, 0);
    for (int k = 0, placed = 0; placed < nem && k < nem * 8; ++k) {
        uint32_t sidx = r.next() % (uint32_t)nsph, tidx = r.next() % (uint32_t)tps;
        uint8_t& e = em[(size_t)sidx * tps + tidx];
        if (!e) { e = 1; placed++; }
    }
    std::vector<std::vector<rrh_vertex>> gv(nmat);
    std::vector<std::vector<uint32_t>> gi(nmat);
    struct EmTri { V3 a, b, c, na, nb, nc; };
    std::vector<EmTri> emt;
    for (int sp = 0; sp < nsph; ++sp) {
        float rad = r.logu(0.05f, 2.0f);
        float cx = r.range(-0.45f * ext, 0.45f * ext), cz = r.range(-0.45f * ext, 0.45f * ext);
        float cy = height(cx, cz) + rad * r.range(0.3f, 1.5f);
        uint32_t m = r.next() % (uint32_t)nmat;
        auto& v = gv[m];
        auto& i = gi[m];
        uint32_t base = (uint32_t)v.size();
        for (const V3& p : P) v.push_back(make_v(v3(cx + rad * p.x, cy + rad * p.y, cz + rad * p.z), p, 0, 0));
        for (int t = 0; t < tps; ++t) {
            uint32_t a = I[3 * t], b = I[3 * t + 1], c = I[3 * t + 2];
            if (em[(size_t)sp * tps + t]) {
                auto W = [&](uint32_t q) { return v3(cx + rad * P[q].x, cy + rad * P[q].y, cz + rad * P[q].z); };
                emt.push_back(EmTri{W(a), W(b), W(c), P[a], P[b], P[c]});
            } else {
                i.push_back(base + a); i.push_back(base + b); i.push_back(base + c);
            }
        }
    }
    for (int m = 0; m < nmat; ++m)
        if (!gi[m].empty()) add_mesh(s, gv[m], gi[m], mat0 + m);
    for (const EmTri& e : emt) {
        float L = r.logu(1.0f, 100.0f);
        float tr = r.range(0.6f, 1.0f), tg = r.range(0.6f, 1.0f), tb = r.range(0.6f, 1.0f);
        s.mats.push_back(material(1, 1, 1, 0.0f, 1.0f, L * tr, L * tg, L * tb));
        std::vector<rrh_vertex> v = {make_v(e.a, e.na, 0, 0), make_v(e.b, e.nb, 1, 0), make_v(e.c, e.nc, 0, 1)};
        std::vector<uint32_t> i = {0, 1, 2};
        add_mesh(s, v, i, (uint32_t)s.mats.size() - 1);
    }
    s.flags = RRH_SCENE_TRI_LIGHTS;
    s.eye[0] = 0.0f; s.eye[1] = height(0.0f, 60.0f) + 30.0f; s.eye[2] = 60.0f;
    s.center[0] = 0.0f; s.center[1] = height(0.0f, 0.0f); s.center[2] = 0.0f;
    s.fov = 45.0f;
}

// closed 10 x 10 x 4 room, tessellated walls, 0.5 x 0.5 window in the x = 0 wall,
// furniture boxes (tables / cabinets with small objects on the floor and on their tops),
// lit by a directional sun through the window.  scale 1 -> about 1.0M triangles.
void indoor(SceneData& s, uint32_t seed, float scale) {
    Rng r(seed);
    const float X = 10.0f, Y = 4.0f, Z = 10.0f;
    int nf = std::max(8, (int)std::lround(370.0f * scale));       // cells along 10 m
    int nh = std::max(4, (int)std::lround(148.0f * scale));       // cells along 4 m
    int nbox = std::max(16, (int)std::lround(1500.0f * scale));
    auto lam = [&](float lo, float hi) {
        s.mats.push_back(material(r.range(lo, hi), r.range(lo, hi), r.range(lo, hi), 0.0f, 1.0f, 0, 0, 0, RRH_MAT_PURE_LAMBERT));
        return (uint32_t)s.mats.size() - 1;
    };
    // window in the x = 0 wall: z in [4.75, 5.25], y in [2.25, 2.75]
    const float wz0 = 4.75f, wz1 = 5.25f, wy0 = 2.25f, wy1 = 2.75f;
    // plane grid: origin o, axes u (nu cells) and v (nv cells), normal n; skip cells inside the window
    auto plane = [&](V3 o, V3 u, V3 v, int nu, int nv, V3 n, uint32_t m, bool window) {
        std::vector<rrh_vertex> vs; std::vector<uint32_t> is;
        const bool flip = (u.y * v.z - u.z * v.y) * n.x + (u.z * v.x - u.x * v.z) * n.y + (u.x * v.y - u.y * v.x) * n.z < 0.0f;
        for (int j = 0; j <= nv; ++j)
            for (int i = 0; i <= nu; ++i) {
                float a = (float)i / nu, b = (float)j / nv;
                vs.push_back(make_v(v3(o.x + u.x * a + v.x * b, o.y + u.y * a + v.y * b, o.z + u.z * a + v.z * b), n, a, b));
            }
        for (int j = 0; j < nv; ++j)
            for (int i = 0; i < nu; ++i) {
                if (window) {
                    const rrh_vertex& c0 = vs[j * (nu + 1) + i];
                    const rrh_vertex& c1 = vs[(j + 1) * (nu + 1) + i + 1];
                    float zc = 0.5f * (c0.pos[2] + c1.pos[2]), yc = 0.5f * (c0.pos[1] + c1.pos[1]);
                    if (zc > wz0 && zc < wz1 && yc > wy0 && yc < wy1) continue;
                }
                uint32_t a = j * (nu + 1) + i, b = a + 1, c = a + (nu + 1), d = c + 1;
                if (flip) std::swap(b, c);  // triangles face the plane's normal
                uint32_t t[6] = {a, b, d, a, d, c};
                is.insert(is.end(), t, t + 6);
            }
        add_mesh(s, vs, is, m);
    };
    plane(v3(0, 0, 0), v3(X, 0, 0), v3(0, 0, Z), nf, nf, v3(0, 1, 0), lam(0.4f, 0.7f), false);      // floor
    plane(v3(0, Y, 0), v3(X, 0, 0), v3(0, 0, Z), nf, nf, v3(0, -1, 0), lam(0.7f, 0.9f), false);     // ceiling
    plane(v3(0, 0, 0), v3(0, 0, Z), v3(0, Y, 0), nf, nh, v3(1, 0, 0), lam(0.5f, 0.8f), true);       // x = 0 (window)
    plane(v3(X, 0, 0), v3(0, 0, Z), v3(0, Y, 0), nf, nh, v3(-1, 0, 0), lam(0.5f, 0.8f), false);     // x = X
    plane(v3(0, 0, 0), v3(X, 0, 0), v3(0, Y, 0), nf, nh, v3(0, 0, 1), lam(0.5f, 0.8f), false);      // z = 0
    plane(v3(0, 0, Z), v3(X, 0, 0), v3(0, Y, 0), nf, nh, v3(0, 0, -1), lam(0.5f, 0.8f), false);     // z = Z
    // furniture: a few large pieces and many small ones, grouped by 8 materials
    const int nm = 8;
    uint32_t m0 = (uint32_t)s.mats.size();
    for (int k = 0; k < nm; ++k) lam(0.1f, 0.9f);
    std::vector<std::vector<rrh_vertex>> gv(nm);
    std::vector<std::vector<uint32_t>> gi(nm);
    struct Top { float x0, z0, x1, z1, y; };
    std::vector<Top> tops;
    for (int b = 0; b < nbox; ++b) {
        bool big = b < 12;
        float sx = big ? r.range(0.6f, 2.0f) : r.range(0.05f, 0.3f);
        float sz = big ? r.range(0.6f, 2.0f) : r.range(0.05f, 0.3f);
        float sy = big ? r.range(0.4f, 1.2f) : r.range(0.05f, 0.4f);
        float x, z, y = 0.0f;
        if (!big && !tops.empty() && r.uf() < 0.3f) {  // on a table top
            const Top& t = tops[r.next() % (uint32_t)tops.size()];
            sx = std::min(sx, t.x1 - t.x0);
            sz = std::min(sz, t.z1 - t.z0);
            x = r.range(t.x0, t.x1 - sx);
            z = r.range(t.z0, t.z1 - sz);
            y = t.y;
        } else {
            x = r.range(0.6f, X - 0.6f - sx);
            z = r.range(0.3f, Z - 0.3f - sz);
        }
        if (big) tops.push_back(Top{x, z, x + sx, z + sz, sy});
        uint32_t m = r.next() % (uint32_t)nm;
        std::vector<rrh_vertex>& v = gv[m];
        std::vector<uint32_t>& i = gi[m];
        std::vector<rrh_vertex> bv; std::vector<uint32_t> bi;
        add_box_facing(bv, bi, v3(x, y, z), v3(x + sx, y + sy, z + sz));
        uint32_t base = (uint32_t)v.size();
        v.insert(v.end(), bv.begin(), bv.end());
        for (uint32_t q : bi) i.push_back(base + q);
    }
    for (int m = 0; m < nm; ++m)
        if (!gi[m].empty()) add_mesh(s, gv[m], gi[m], m0 + m);
    rrh_light sun;
    std::memset(&sun, 0, sizeof(sun));
    sun.type = 1;
    sun.direction[0] = 0.8f; sun.direction[1] = -0.45f; sun.direction[2] = 0.12f;  // travels into the room
    sun.emission[0] = 40.0f; sun.emission[1] = 36.0f; sun.emission[2] = 30.0f;
    sun.intensity = 1.0f;
    s.lights.push_back(sun);
    s.flags = 0;
    s.eye[0] = 8.5f; s.eye[1] = 1.6f; s.eye[2] = 8.5f;
    s.center[0] = 2.0f; s.center[1] = 1.0f; s.center[2] = 4.0f;
    s.fov = 60.0f;
}

}  // namespace

bool generate_scene(int kind, uint32_t seed, float scale, SceneData& out, std::string& err) {
    out = SceneData();
    if (!(scale > 0.0f && scale <= 4.0f)) { err = "scale must be in (0, 4]"; return false; }
    switch (kind) {
    case RRH_GEN_CORNELL_REF: cornell_ref(out); return true;
    case RRH_GEN_CORNELL_RANDOM: cornell_random(out, seed); return true;
    case RRH_GEN_HEIGHTFIELD: heightfield(out, seed, scale); return true;
    case RRH_GEN_INDOOR: indoor(out, seed, scale); return true;
    default: err = "unknown scene kind"; return false;
    }
}

}  // namespace rrh

template <class T>
static T* dup(const std::vector<T>& v) {
    T* p = (T*)std::malloc(v.size() * sizeof(T) + 16);
    if (p && !v.empty()) std::memcpy(p, v.data(), v.size() * sizeof(T));
    return p;
}

extern "C" int rrh_scene_generate(int kind, uint32_t seed, float scale, rrh_scene_desc* out) {
    if (!out) { rrh::set_error("rrh_scene_generate: out is null"); return RRH_E_ARG; }
    std::memset(out, 0, sizeof(*out));
    rrh::SceneData s;
    std::string err;
    if (!rrh::generate_scene(kind, seed, scale, s, err)) { rrh::set_error("rrh_scene_generate: " + err); return RRH_E_ARG; }
    out->verts = dup(s.verts); out->nv = (uint32_t)s.verts.size();
    out->idx = dup(s.idx); out->ni = (uint32_t)s.idx.size();
    out->meshes = dup(s.meshes); out->nm = (uint32_t)s.meshes.size();
    out->mats = dup(s.mats); out->nmat = (uint32_t)s.mats.size();
    out->insts = dup(s.insts); out->ninst = (uint32_t)s.insts.size();
    out->lights = dup(s.lights); out->nl = (uint32_t)s.lights.size();
    out->flags = s.flags;
    for (int k = 0; k < 3; ++k) { out->cam_eye[k] = s.eye[k]; out->cam_center[k] = s.center[k]; }
    out->fov_y_deg = s.fov;
    if (!out->verts || !out->idx || !out->meshes || !out->mats || !out->insts || !out->lights) {
        rrh_scene_desc_free(out);
        rrh::set_error("rrh_scene_generate: out of host memory");
        return RRH_E_NOMEM;
    }
    return RRH_OK;
}

extern "C" int rrh_scene_desc_free(rrh_scene_desc* d) {
    if (!d) return RRH_OK;
    std::free(d->verts); std::free(d->idx); std::free(d->meshes); std::free(d->mats); std::free(d->insts); std::free(d->lights);
    std::memset(d, 0, sizeof(*d));
    return RRH_OK;
}

// render.cpp -- device, images, camera, scene and renderer of the host layer (rr_host.h).
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "rr_host.h"

namespace rrh {

void check(int rc, const char* what) {
    if (rc != RRH_OK) throw std::runtime_error(std::string(what) + ": " + rrh_last_error());
}

namespace rhi {

void Device::create(int device) {
    if (ctx_) return;
    check(rrh_ctx_create(device, &ctx_), "rrh_ctx_create");
}
void Device::destroy() {
    if (ctx_) rrh_ctx_destroy(ctx_);
    ctx_ = nullptr;
}
void Device::wait_idle() {
    if (ctx_) check(rrh_ctx_synchronize(ctx_), "rrh_ctx_synchronize");
}

static int to_rrh(Format f) {
    switch (f) {
        case Format::RGBA32F: return RRH_FMT_RGBA32F;
        case Format::RGBA16F: return RRH_FMT_RGBA16F;
        case Format::RGBA64F: return RRH_FMT_RGBA64F;
        case Format::R32U: return RRH_FMT_R32U;
        case Format::RGBA8: return RRH_FMT_RGBA8;
        default: throw std::runtime_error("image format undefined");
    }
}
static size_t texel_bytes(Format f) {
    switch (f) {
        case Format::RGBA32F: return 16;
        case Format::RGBA16F: return 8;
        case Format::RGBA64F: return 32;
        case Format::R32U: return 4;
        case Format::RGBA8: return 4;
        default: return 0;
    }
}

void Image::create(Device& d, Extent2D e, Format f) {
    destroy();
    check(rrh_image_create(d.ctx(), e.width, e.height, to_rrh(f), &img_), "rrh_image_create");
    extent_ = e;
    format_ = f;
    check(rrh_image_clear(d.ctx(), img_), "rrh_image_clear");
}
void Image::destroy() {
    if (img_) rrh_image_destroy(img_);
    img_ = nullptr;
}
void Image::clear(Device& d) { check(rrh_image_clear(d.ctx(), img_), "rrh_image_clear"); }
size_t Image::bytes() const { return (size_t)extent_.width * extent_.height * texel_bytes(format_); }
void Image::download(Device& d, void* host, size_t bytes) const {
    check(rrh_image_download(d.ctx(), img_, host, bytes), "rrh_image_download");
}
void Image::upload(Device& d, const void* host, size_t bytes) {
    check(rrh_image_upload(d.ctx(), img_, host, bytes), "rrh_image_upload");
}

}  // namespace rhi

namespace scene {

void Camera::look_at(std::array<float, 3> eye, std::array<float, 3> center, std::array<float, 3> /*up*/) {
    position_ = eye;
    pivot_ = center;
    double d[3] = {(double)center[0] - eye[0], (double)center[1] - eye[1], (double)center[2] - eye[2]};
    double len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (len > 0.0) {
        pitch_ = (float)(std::asin(d[1] / len) * 180.0 / M_PI);
        yaw_ = (float)(std::atan2(d[2], d[0]) * 180.0 / M_PI);
    }
}
std::array<float, 3> Camera::forward() const {
    const double y = yaw_ * M_PI / 180.0, p = pitch_ * M_PI / 180.0;
    double f[3] = {std::cos(y) * std::cos(p), std::sin(p), std::sin(y) * std::cos(p)};
    double len = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    return {(float)(f[0] / len), (float)(f[1] / len), (float)(f[2] / len)};
}

void Scene::load_generated(rhi::Device& d, int kind, uint32_t seed, float scale) {
    rrh_scene_desc desc{};
    check(rrh_scene_generate(kind, seed, scale, &desc), "rrh_scene_generate");
    static const char* names[] = {"cornell_box", "cornell_random", "heightfield_icospheres", "indoor_room"};
    try {
        upload(d, desc, (kind >= 0 && kind < 4) ? names[kind] : "generated");
    } catch (...) {
        rrh_scene_desc_free(&desc);
        throw;
    }
    rrh_scene_desc_free(&desc);
}

void Scene::upload(rhi::Device& d, const rrh_scene_desc& s, const std::string& name) {
    rrh_scene* out = nullptr;
    check(rrh_scene_upload(d.ctx(), s.verts, s.nv, s.idx, s.ni, s.meshes, s.nm, s.mats, s.nmat, s.insts, s.ninst,
                           s.nl ? s.lights : nullptr, s.nl, s.flags, &out),
          "rrh_scene_upload");
    destroy();
    scene_ = out;
    name_ = name;
    for (int k = 0; k < 3; ++k) {
        suggested_eye[k] = s.cam_eye[k];
        suggested_center[k] = s.cam_center[k];
    }
    suggested_fov = s.fov_y_deg;
}

void Scene::upload(rhi::Device& d, const GltfScene& g, const std::string& name) {
    rrh_scene_desc desc{};
    desc.verts = const_cast<rrh_vertex*>(g.vertices.data());
    desc.nv = (uint32_t)g.vertices.size();
    desc.idx = const_cast<uint32_t*>(g.indices.data());
    desc.ni = (uint32_t)g.indices.size();
    desc.meshes = const_cast<rrh_mesh*>(g.meshes.data());
    desc.nm = (uint32_t)g.meshes.size();
    desc.mats = const_cast<rrh_material*>(g.materials.data());
    desc.nmat = (uint32_t)g.materials.size();
    desc.insts = const_cast<rrh_instance*>(g.instances.data());
    desc.ninst = (uint32_t)g.instances.size();
    desc.cam_eye[2] = 3.0f;  // the default Camera (Camera.h:55-64)
    desc.fov_y_deg = 45.0f;
    upload(d, desc, name);
}

void Scene::destroy() {
    if (scene_) rrh_scene_destroy(scene_);
    scene_ = nullptr;
    have_view_proj_ = false;
}

rrh_scene_info Scene::info() const {
    rrh_scene_info i{};
    check(rrh_scene_info_get(scene_, &i), "rrh_scene_info_get");
    return i;
}

void Scene::update_camera(const Camera& cam, uint32_t width, uint32_t height, uint32_t frame_index) {
    // the view looks along the camera's forward from its position (Camera::rebuild_matrices)
    const auto p = cam.position();
    const auto f = cam.forward();
    const float eye[3] = {p[0], p[1], p[2]};
    const float ctr[3] = {p[0] + f[0], p[1] + f[1], p[2] + f[2]};
    rrh_camera c{};
    check(rrh_camera_look_at(eye, ctr, cam.fov_y_deg, width, height, cam.near_plane, cam.far_plane, frame_index, &c),
          "rrh_camera_look_at");
    // rrh_camera_look_at fills prev_view_proj with this view's matrix; the previous call's
    // takes its place, and this one is kept for the next frame
    float vp[16];
    std::memcpy(vp, c.prev_view_proj, sizeof(vp));
    if (have_view_proj_) std::memcpy(c.prev_view_proj, last_view_proj_, sizeof(vp));
    std::memcpy(last_view_proj_, vp, sizeof(vp));
    have_view_proj_ = true;
    check(rrh_camera_set(scene_, &c), "rrh_camera_set");
}

}  // namespace scene

namespace render {

void Renderer::render(FrameContext& fc) {
    fc.renderer = this;
    for (auto& p : passes_)
        if (p->is_enabled()) p->execute(fc);
}
void Renderer::on_resize(rhi::Extent2D e) {
    for (auto& p : passes_) p->on_resize(e);
}
RenderPass* Renderer::find(const std::string& name) const {
    for (auto& p : passes_)
        if (name == p->name()) return p.get();
    return nullptr;
}

}  // namespace render
}  // namespace rrh

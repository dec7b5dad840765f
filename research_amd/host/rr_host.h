// rr_host.h -- the host side of the renderer above the C ABI (include/rrh.h): the render-pass
// interface, the passes of the path (path tracer, accumulation, ReSTIR DI, A-trous, guiding),
// the renderer that runs them in order, and the headless Application.
//
// Names, knobs and defaults follow the reference's classes so code written against them
// ports by changing the namespace:
//   RenderPass            src/render/RenderPass.h:11-48
//   FrameContext          src/render/FrameContext.h:26-41
//   Renderer::add_pass    src/render/Renderer.h:18-25, RenderGraph::execute RenderGraph.cpp:14-21
//   PathTracerPass        src/passes/pathtracer_reference/PathTracerPass.h:20-75
//   AccumulatePass        src/passes/accumulate/AccumulatePass.h:20-70
//   ReSTIRDIPass          src/passes/restir_di/ReSTIRDIPass.h:26-136
//   AtrousPass            src/passes/denoise/AtrousPass.h:16-90
//   GuidingPass           the guiding knobs of src/passes/restir_pt/ReSTIRPTPass.h:93-100
//   Camera / Scene        src/scene/Camera.cpp:27-36, Scene.cpp:144-251,554-585
//   Application           src/app/Application.h:38-100, bindings src/python/bindings.cpp:28-100
// Device objects are created by initialize() and released by shutdown(); errors from the C ABI
// surface as std::runtime_error carrying rrh_last_error().
#pragma once

#include <array>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "rrh.h"

namespace rrh {

// throws std::runtime_error(what + rrh_last_error()) when rc != RRH_OK
void check(int rc, const char* what);

namespace rhi {
struct Extent2D {
    uint32_t width = 0, height = 0;
};
enum class Format { Undefined, RGBA32F, RGBA16F, RGBA64F, R32U, RGBA8 };

// One HIP device and its stream (replaces rr::rhi::Device).
class Device {
public:
    void create(int device = 0);
    void destroy();
    void wait_idle();
    [[nodiscard]] rrh_ctx* ctx() const { return ctx_; }
    [[nodiscard]] bool valid() const { return ctx_ != nullptr; }

private:
    rrh_ctx* ctx_ = nullptr;
};

// A device image (replaces rr::rhi::Image): explicit create / destroy like the reference.
class Image {
public:
    void create(Device& d, Extent2D e, Format f);
    void destroy();
    void clear(Device& d);
    void download(Device& d, void* host, size_t bytes) const;
    void upload(Device& d, const void* host, size_t bytes);
    [[nodiscard]] rrh_image* handle() const { return img_; }
    [[nodiscard]] Extent2D extent() const { return extent_; }
    [[nodiscard]] Format format() const { return format_; }
    [[nodiscard]] size_t bytes() const;

private:
    rrh_image* img_ = nullptr;
    Extent2D extent_{};
    Format format_ = Format::Undefined;
};
}  // namespace rhi

namespace scene {
// Fly camera state; look_at keeps position, pivot and yaw / pitch like Camera::look_at.
class Camera {
public:
    void look_at(std::array<float, 3> eye, std::array<float, 3> center, std::array<float, 3> up = {0, 1, 0});
    [[nodiscard]] std::array<float, 3> position() const { return position_; }
    [[nodiscard]] std::array<float, 3> forward() const;
    [[nodiscard]] std::array<float, 3> pivot() const { return pivot_; }
    float fov_y_deg = 45.0f, near_plane = 0.01f, far_plane = 1000.0f;

private:
    std::array<float, 3> position_{0.0f, 0.0f, 3.0f};
    std::array<float, 3> pivot_{0.0f, 0.0f, 0.0f};
    float yaw_ = -90.0f, pitch_ = 0.0f;
};

// glTF 2.0 import (GltfLoader::load, src/scene/GltfLoader.cpp:79-363): flat records ready for
// rrh_scene_upload.  Images are described, not decoded (no kernel samples textures).
struct GltfImage {
    std::string name, mime_type, uri;
    uint32_t width = 0, height = 0;  // from the PNG / JPEG header; 0 = unusable
    size_t bytes = 0;
};
struct GltfMaterialImages {  // image index per texture slot, -1 = none
    int32_t base_color = -1, metallic_roughness = -1, normal = -1, emissive = -1;
};
struct GltfScene {
    std::string name;  // file stem
    std::vector<rrh_vertex> vertices;
    std::vector<uint32_t> indices;
    std::vector<rrh_mesh> meshes;
    std::vector<std::string> mesh_names;
    std::vector<rrh_material> materials;
    std::vector<std::string> material_names;
    std::vector<GltfMaterialImages> material_images;
    std::vector<rrh_instance> instances;  // row-major world transforms
    std::vector<GltfImage> images;
    std::vector<std::string> warnings;
};
// ".glb" -> binary container, otherwise JSON; throws std::runtime_error on malformed input
void load_gltf(const std::string& path, GltfScene& out);
std::string gltf_stem(const std::string& path);

// Uploaded scene (flattened triangles + BVH + shading records on the device).
class Scene {
public:
    // procedural scenes: RRH_GEN_*; the reference Cornell box is RRH_GEN_CORNELL_REF
    void load_generated(rhi::Device& d, int kind, uint32_t seed = 1, float scale = 1.0f);
    void upload(rhi::Device& d, const rrh_scene_desc& desc, const std::string& name);
    // Scene records from a glTF file (no lights, no scene flags: GltfLoader adds neither)
    void upload(rhi::Device& d, const GltfScene& g, const std::string& name);
    void destroy();
    // Scene::update_camera: camera matrices for the given extent and frame; prev_view_proj is
    // the view-projection of the previous call (Camera::update keeps it, Camera.cpp:66)
    void update_camera(const Camera& cam, uint32_t width, uint32_t height, uint32_t frame_index);
    [[nodiscard]] rrh_scene* handle() const { return scene_; }
    [[nodiscard]] const std::string& name() const { return name_; }
    [[nodiscard]] rrh_scene_info info() const;
    [[nodiscard]] bool valid() const { return scene_ != nullptr; }
    // camera the generator suggests for this scene
    std::array<float, 3> suggested_eye{0, 0, 3}, suggested_center{0, 0, 0};
    float suggested_fov = 45.0f;

private:
    rrh_scene* scene_ = nullptr;
    std::string name_;
    float last_view_proj_[16] = {};
    bool have_view_proj_ = false;
};
}  // namespace scene

namespace render {
class Renderer;

struct FrameContext {
    rhi::Device* device = nullptr;
    scene::Scene* scene = nullptr;
    Renderer* renderer = nullptr;
    rhi::Extent2D extent{};
    uint32_t frame_index = 0;
    float delta_time_seconds = 0.0f;
    const rhi::Image* accumulated_image = nullptr;  // what the display (screenshot) shows
};

class RenderPass {
public:
    struct ResourceDesc {
        enum class Kind { Texture, Buffer };
        std::string name;
        Kind kind = Kind::Texture;
        rhi::Format format = rhi::Format::Undefined;
        rhi::Extent2D extent{};
        bool persistent = false;
    };
    struct Reflection {
        std::vector<ResourceDesc> inputs;
        std::vector<ResourceDesc> outputs;
    };
    virtual ~RenderPass() = default;
    [[nodiscard]] virtual const char* name() const = 0;
    [[nodiscard]] virtual Reflection reflect() const = 0;
    virtual void render_ui() {}
    virtual void on_resize(rhi::Extent2D /*new_extent*/) {}
    virtual void execute(FrameContext& fc) = 0;
    void set_enabled(bool v) { enabled_ = v; }
    [[nodiscard]] bool is_enabled() const { return enabled_; }

protected:
    bool enabled_ = true;
};

// Owns the passes and runs the enabled ones in insertion order (RenderGraph::execute).
class Renderer {
public:
    template <class T>
    T* add_pass() {
        passes_.push_back(std::make_unique<T>());
        return static_cast<T*>(passes_.back().get());
    }
    void render(FrameContext& fc);
    void on_resize(rhi::Extent2D e);
    [[nodiscard]] RenderPass* find(const std::string& name) const;
    [[nodiscard]] const std::vector<std::unique_ptr<RenderPass>>& passes() const { return passes_; }
    void clear() { passes_.clear(); }

private:
    std::vector<std::unique_ptr<RenderPass>> passes_;
};
}  // namespace render

namespace passes {
using render::FrameContext;
using render::RenderPass;

// G-buffer (GBufferPass.h:19-90): position RGBA32F (xyz, 1), normal RGBA32F (xyz,
// asfloat(material)), material id R32U, from pixel-centre rays with back faces culled.
class GBufferPass : public RenderPass {
public:
    ~GBufferPass() override;
    void initialize(rhi::Device& device, rhi::Extent2D extent);
    void shutdown(rhi::Device& device);
    bool reload_shader() { return true; }
    [[nodiscard]] const char* name() const override { return "GBufferPass"; }
    [[nodiscard]] Reflection reflect() const override;
    void on_resize(rhi::Extent2D new_extent) override;
    void execute(FrameContext& fc) override;
    [[nodiscard]] const rhi::Image& position_image() const { return position_img_; }
    [[nodiscard]] const rhi::Image& normal_image() const { return normal_img_; }
    [[nodiscard]] const rhi::Image& material_id_image() const { return material_id_img_; }
    [[nodiscard]] rhi::Device* device() const { return device_; }

private:
    rhi::Device* device_ = nullptr;
    rhi::Image position_img_, normal_img_, material_id_img_;
    rhi::Extent2D extent_{};
    bool initialized_ = false;
};

// Tonemap (TonemapPass / tonemap.slang, the screenshot's Application.cpp:853-866): ACES +
// gamma 1/2.2 into an RGBA8 image, on the device.
class TonemapPass : public RenderPass {
public:
    ~TonemapPass() override;
    void initialize(rhi::Device& device, rhi::Extent2D extent);
    void shutdown(rhi::Device& device);
    bool reload_shader() { return true; }
    [[nodiscard]] const char* name() const override { return "TonemapPass"; }
    [[nodiscard]] Reflection reflect() const override;
    void on_resize(rhi::Extent2D new_extent) override;
    void execute(FrameContext& fc) override;
    void set_input(const rhi::Image* hdr) { input_ = hdr; }
    // tonemap `hdr` now (the display path calls this with the current exposure)
    void run(const rhi::Image& hdr);
    [[nodiscard]] const rhi::Image& output_image() const { return ldr_img_; }

    float exposure = 1.0f;

private:
    rhi::Device* device_ = nullptr;
    const rhi::Image* input_ = nullptr;
    rhi::Image ldr_img_;
    rhi::Extent2D extent_{};
    bool initialized_ = false;
};

// Reference path tracer: one sample per pixel per frame into an RGBA32F radiance image.
class PathTracerPass : public RenderPass {
public:
    ~PathTracerPass() override;
    void initialize(rhi::Device& device, rhi::Extent2D extent);
    void shutdown(rhi::Device& device);
    bool reload_shader() { return true; }  // kernels are compiled ahead of time
    [[nodiscard]] const char* name() const override { return "PathTracerPass"; }
    [[nodiscard]] Reflection reflect() const override;
    void on_resize(rhi::Extent2D new_extent) override;
    void execute(FrameContext& fc) override;
    [[nodiscard]] const rhi::Image& output_image() const { return radiance_img_; }
    [[nodiscard]] uint32_t frame_count() const { return frame_count_; }

    uint32_t max_bounces = 5;
    uint32_t rr_start = 3;  // pathtracer.slang:252

private:
    rhi::Device* device_ = nullptr;
    rhi::Image radiance_img_;
    rhi::Extent2D extent_{};
    uint32_t frame_count_ = 0;
    bool initialized_ = false;
};

// Progressive running mean (accumulate.slang); optional FP64 running sum beside it.
class AccumulatePass : public RenderPass {
public:
    ~AccumulatePass() override;
    void initialize(rhi::Device& device, rhi::Extent2D extent);
    void shutdown(rhi::Device& device);
    bool reload_shader() { return true; }
    [[nodiscard]] const char* name() const override { return "AccumulatePass"; }
    [[nodiscard]] Reflection reflect() const override;
    void on_resize(rhi::Extent2D new_extent) override;
    void execute(FrameContext& fc) override;
    void set_input(const rhi::Image* radiance) { input_ = radiance; }
    [[nodiscard]] const rhi::Image& accumulated_image() const { return accum_img_; }
    [[nodiscard]] const rhi::Image& accumulated_f64_image() const { return accum64_img_; }

    bool camera_moved = false;
    uint32_t accumulated_spp = 0;
    bool fp64_enable = false;

private:
    rhi::Device* device_ = nullptr;
    const rhi::Image* input_ = nullptr;
    rhi::Image accum_img_, accum64_img_;
    rhi::Extent2D extent_{};
    bool initialized_ = false;
};

// ReSTIR DI: initial candidates, temporal and spatial reuse, shading.  Writes its own
// primary-hit G-buffer (positions, normals) which the A-trous pass reads.
class ReSTIRDIPass : public RenderPass {
public:
    ~ReSTIRDIPass() override;
    void initialize(rhi::Device& device, rhi::Extent2D extent);
    void shutdown(rhi::Device& device);
    bool reload_shader() { return true; }
    [[nodiscard]] const char* name() const override { return "ReSTIRDIPass"; }
    [[nodiscard]] Reflection reflect() const override;
    void on_resize(rhi::Extent2D new_extent) override;
    void execute(FrameContext& fc) override;
    void reset_history() {
        reservoir_flip_ = 0;
        first_execute_ = true;
        frame_count_ = 0;
    }
    [[nodiscard]] const rhi::Image& output_image() const { return output_img_; }
    [[nodiscard]] const rhi::Image& gbuffer_position() const { return gbuf_pos_; }
    [[nodiscard]] const rhi::Image& gbuffer_normal() const { return gbuf_norm_; }

    bool temporal_enable = true;
    bool spatial_enable = true;
    uint32_t num_initial_candidates = 16;
    uint32_t spatial_num_neighbors = 4;
    uint32_t spatial_radius = 15;

private:
    rhi::Device* device_ = nullptr;
    rhi::Image gbuf_pos_, gbuf_norm_, output_img_;
    rhi::Image reservoir_[2], reservoir_uv_[2];
    uint32_t reservoir_flip_ = 0;
    rhi::Extent2D extent_{};
    uint32_t frame_count_ = 0;
    bool first_execute_ = true;
    bool initialized_ = false;
};

// ReSTIR PT with the reconnection shift (ReSTIRPTPass.h:22-165): initial multi-bounce
// candidates, temporal and spatial reuse, shading over ReSTIR DI's G-buffer and direct image,
// first-bounce guiding by the reference's single-lobe hash grid (splat, then fit each frame).
class ReSTIRPTPass : public RenderPass {
public:
    ~ReSTIRPTPass() override;
    void initialize(rhi::Device& device, rhi::Extent2D extent);
    void shutdown(rhi::Device& device);
    bool reload_shader() { return true; }
    [[nodiscard]] const char* name() const override { return "ReSTIRPTPass"; }
    [[nodiscard]] Reflection reflect() const override;
    void on_resize(rhi::Extent2D new_extent) override;
    void execute(FrameContext& fc) override;
    // G-buffer position / normal (ReSTIR DI's) and the direct lighting image
    void set_inputs(const rhi::Image* gbuf_pos, const rhi::Image* gbuf_norm, const rhi::Image* direct) {
        gbuf_pos_ = gbuf_pos;
        gbuf_norm_ = gbuf_norm;
        direct_ = direct;
    }
    void reset_history() {
        reservoir_flip_ = 0;
        frame_count_ = 0;
        history_dirty_ = true;
        grid_dirty_ = true;
    }
    [[nodiscard]] const rhi::Image& output_image() const { return output_img_; }
    [[nodiscard]] rrh_grid* guiding_grid() const { return grid_; }

    bool include_direct_lighting = true;
    bool temporal_enable = true;
    bool spatial_enable = true;
    float indirect_strength = 1.0f;
    float ray_bias = 1.0e-3f;
    uint32_t num_initial_candidates = 8;
    uint32_t max_bounces = 4;
    uint32_t spatial_num_neighbors = 4;
    uint32_t spatial_radius = 12;
    bool guiding_enable = true;
    float guiding_alpha = 0.5f;
    float grid_base_cell_size = 0.10f;
    float grid_distance_slope = 0.02f;
    uint32_t grid_min_samples = 8;
    float grid_kappa_max = 512.0f;
    static constexpr uint32_t kGridCellCount = 1u << 20;

private:
    void destroy_images();
    rhi::Device* device_ = nullptr;
    const rhi::Image *gbuf_pos_ = nullptr, *gbuf_norm_ = nullptr, *direct_ = nullptr;
    rhi::Image reservoir_[2][4];
    rhi::Image history_pos_[2], history_norm_[2];
    rhi::Image output_img_;
    rrh_grid* grid_ = nullptr;
    float grid_made_[3] = {0, 0, 0};  // base cell, slope, kappa of the grid as created
    uint32_t grid_made_min_ = 0;
    rhi::Extent2D extent_{};
    bool initialized_ = false, history_dirty_ = true, grid_dirty_ = true;
    uint32_t reservoir_flip_ = 0, frame_count_ = 0;
};

// ReSTIR GI (ReSTIRGIPass.h:22-122): one-bounce indirect reuse with the reconnection shift over
// ReSTIR DI's G-buffer and direct image; cosine candidates, temporal and spatial reuse.
class ReSTIRGIPass : public RenderPass {
public:
    ~ReSTIRGIPass() override;
    void initialize(rhi::Device& device, rhi::Extent2D extent);
    void shutdown(rhi::Device& device);
    bool reload_shader() { return true; }
    [[nodiscard]] const char* name() const override { return "ReSTIRGIPass"; }
    [[nodiscard]] Reflection reflect() const override;
    void on_resize(rhi::Extent2D new_extent) override;
    void execute(FrameContext& fc) override;
    void set_inputs(const rhi::Image* gbuf_pos, const rhi::Image* gbuf_norm, const rhi::Image* direct) {
        gbuf_pos_ = gbuf_pos;
        gbuf_norm_ = gbuf_norm;
        direct_ = direct;
    }
    void reset_history() {
        reservoir_flip_ = 0;
        frame_count_ = 0;
        history_dirty_ = true;
    }
    [[nodiscard]] const rhi::Image& output_image() const { return output_img_; }

    bool include_direct_lighting = true;
    bool temporal_enable = true;
    bool spatial_enable = true;
    float indirect_strength = 1.0f;
    float ray_bias = 1.0e-3f;
    uint32_t num_initial_candidates = 8;
    uint32_t spatial_num_neighbors = 4;
    uint32_t spatial_radius = 12;

private:
    void destroy_images();
    rhi::Device* device_ = nullptr;
    const rhi::Image *gbuf_pos_ = nullptr, *gbuf_norm_ = nullptr, *direct_ = nullptr;
    rhi::Image reservoir_[2][3];
    rhi::Image history_pos_[2], history_norm_[2];
    rhi::Image output_img_;
    rhi::Extent2D extent_{};
    bool initialized_ = false, history_dirty_ = true;
    uint32_t reservoir_flip_ = 0, frame_count_ = 0;
};

// Edge-aware A-trous wavelet filter over the ReSTIR DI output and its G-buffer.
class AtrousPass : public RenderPass {
public:
    ~AtrousPass() override;
    void initialize(rhi::Device& device, rhi::Extent2D extent);
    void shutdown(rhi::Device& device);
    bool reload_shader() { return true; }
    [[nodiscard]] const char* name() const override { return "AtrousPass"; }
    [[nodiscard]] Reflection reflect() const override;
    void on_resize(rhi::Extent2D new_extent) override;
    void execute(FrameContext& fc) override;
    void set_input(const rhi::Image* input) { input_ = input; }
    void set_gbuffer(const rhi::Image* pos, const rhi::Image* norm) {
        gbuf_pos_ = pos;
        gbuf_norm_ = norm;
    }
    [[nodiscard]] const rhi::Image& output_image() const { return ping_images_[out_slot_ < 0 ? 0 : out_slot_]; }

    uint32_t iterations = 5;
    float sigma_position = 0.25f;
    float sigma_normal = 0.20f;
    float sigma_luminance = 1.00f;
    bool fast_math = false;  // hardware exp / reciprocal in the weights (within 1e-3 of exact)

private:
    rhi::Device* device_ = nullptr;
    const rhi::Image* input_ = nullptr;
    const rhi::Image *gbuf_pos_ = nullptr, *gbuf_norm_ = nullptr;
    rhi::Image ping_images_[2];
    int out_slot_ = -1;
    rhi::Extent2D extent_{};
    bool initialized_ = false;
};

// Guided path tracing.  mode "grid": the reference hash grid's knobs apply to the 8-lobe EM
// grid's fallback; mode "vmm": 8-lobe vMF grid fitted by online EM; mode "mlp": the learned
// guide.  Each frame renders one guided sample per pixel, records path vertices and trains.
class GuidingPass : public RenderPass {
public:
    enum class Mode { VMM, MLP };
    ~GuidingPass() override;
    void initialize(rhi::Device& device, rhi::Extent2D extent);
    void shutdown(rhi::Device& device);
    bool reload_shader() { return true; }
    [[nodiscard]] const char* name() const override { return "GuidingPass"; }
    [[nodiscard]] Reflection reflect() const override;
    void on_resize(rhi::Extent2D new_extent) override;
    void execute(FrameContext& fc) override;
    void reset_history();
    [[nodiscard]] const rhi::Image& output_image() const { return radiance_img_; }
    [[nodiscard]] float last_loss() const { return last_loss_; }

    Mode mode = Mode::VMM;
    bool guiding_enable = true;
    float guiding_alpha = 0.5f;     // ReSTIRPTPass.h:95
    uint32_t max_bounces = 12;
    uint32_t grid_resolution = 32;  // VMM cells per axis
    uint32_t grid_lobes = 8;
    float grid_kappa_max = 512.0f;  // ReSTIRPTPass.h:99
    float grid_decay = 0.95f;
    uint32_t record_vertices = 2000000;
    uint32_t mlp_batch = 262144;
    float mlp_learning_rate = 1e-3f;
    uint32_t mlp_steps_per_frame = 1;

private:
    void create_guide(scene::Scene& s);
    void destroy_guide();
    rhi::Device* device_ = nullptr;
    rhi::Image radiance_img_;
    rhi::Extent2D extent_{};
    rrh_grid* grid_ = nullptr;
    rrh_mlp* mlp_ = nullptr;
    const rrh_scene* guide_scene_ = nullptr;
    Mode guide_mode_ = Mode::VMM;
    uint32_t frame_count_ = 0;
    float last_loss_ = 0.0f;
    bool initialized_ = false;
};
}  // namespace passes

namespace app {
// Headless Application: same API as the reference's window-driven one; there is no window,
// so should_close() stays false until request_close() or the frame budget ends the loop.
class Application {
public:
    Application();
    Application(const Application&) = delete;
    Application& operator=(const Application&) = delete;
    ~Application();

    void initialize();
    void shutdown();
    void step();
    [[nodiscard]] bool should_close() const;
    void run_until_closed();

    void load_cornell();
    void load_gltf(const std::string& path);
    void load_procedural(int kind, uint32_t seed, float scale);
    [[nodiscard]] std::string scene_name() const { return current_scene_name_; }
    // why the last glTF request fell back to the Cornell box ("" when it loaded)
    [[nodiscard]] std::string last_scene_error() const { return last_scene_error_; }

    void set_camera_look_at(float ex, float ey, float ez, float cx, float cy, float cz, float ux, float uy, float uz);
    [[nodiscard]] std::array<float, 3> camera_position() const;
    [[nodiscard]] std::array<float, 3> camera_forward() const;

    void set_pass_enabled(const std::string& name, bool enabled);

    void set_exposure(float exposure) { exposure_ = exposure; }
    [[nodiscard]] float exposure() const { return exposure_; }

    void set_input_enabled(bool enabled) { input_enabled_ = enabled; }
    [[nodiscard]] bool input_enabled() const { return input_enabled_; }

    void capture_screenshot(const std::string& path);
    [[nodiscard]] std::array<int, 2> window_size() const { return {width_, height_}; }

    // headless additions
    void set_window_size(int width, int height);
    void request_close() { close_requested_ = true; }
    uint32_t max_frames = 0;  // run_until_closed stops after this many frames (0: needs request_close)
    [[nodiscard]] uint32_t accumulated_spp() const { return accumulated_spp_; }
    [[nodiscard]] uint32_t frames_rendered() const { return frames_rendered_; }
    // the displayed HDR image (RGBA32F, row-major) and its tonemapped LDR form (RGBA8)
    [[nodiscard]] std::vector<float> display_image() const;
    [[nodiscard]] std::vector<uint8_t> display_ldr() const;
    [[nodiscard]] render::Renderer& renderer() { return renderer_; }
    passes::PathTracerPass* pathtracer_pass() { return pathtracer_pass_; }
    passes::AccumulatePass* accumulate_pass() { return accumulate_pass_; }
    passes::ReSTIRDIPass* restir_di_pass() { return restir_di_pass_; }
    passes::AtrousPass* atrous_pass() { return atrous_pass_; }
    passes::GBufferPass* gbuffer_pass() { return gbuffer_pass_; }
    passes::TonemapPass* tonemap_pass() { return tonemap_pass_; }
    passes::GuidingPass* guiding_pass() { return guiding_pass_; }
    passes::ReSTIRPTPass* restir_pt_pass() { return restir_pt_pass_; }
    passes::ReSTIRGIPass* restir_gi_pass() { return restir_gi_pass_; }

private:
    void initialize_renderer();
    void apply_scene_request();
    [[nodiscard]] const rhi::Image* current_display_image() const;

    rhi::Device device_;
    render::Renderer renderer_;
    scene::Scene scene_;
    scene::Camera camera_;
    passes::PathTracerPass* pathtracer_pass_ = nullptr;
    passes::AccumulatePass* accumulate_pass_ = nullptr;
    passes::ReSTIRDIPass* restir_di_pass_ = nullptr;
    passes::AtrousPass* atrous_pass_ = nullptr;
    passes::GBufferPass* gbuffer_pass_ = nullptr;
    passes::TonemapPass* tonemap_pass_ = nullptr;
    passes::GuidingPass* guiding_pass_ = nullptr;
    passes::ReSTIRPTPass* restir_pt_pass_ = nullptr;
    passes::ReSTIRGIPass* restir_gi_pass_ = nullptr;
    const rhi::Image* display_ = nullptr;

    int width_ = 1600, height_ = 900;  // Application.h:158-159
    float exposure_ = 1.0f;
    bool input_enabled_ = true, initialized_ = false, close_requested_ = false;
    bool use_di_ = false, use_denoise_ = false, use_guiding_ = false, use_pt_ = false, use_gi_ = false;
    bool prev_use_di_ = false, prev_use_denoise_ = false, prev_use_guiding_ = false, prev_use_pt_ = false;
    bool prev_use_gi_ = false;
    bool camera_moved_ = true;
    uint32_t accumulated_spp_ = 0, frames_rendered_ = 0;
    std::string current_scene_name_;
    // deferred scene load (applied at the next step, Application.cpp:957-969)
    int pending_kind_ = -1;
    std::string pending_gltf_, last_scene_error_;
    uint32_t pending_seed_ = 1;
    float pending_scale_ = 1.0f;
    static bool s_instance_alive_;
};

// ACES filmic approximation (Narkowicz) + 1/2.2 gamma, as the reference's screenshot path
// (Application.cpp:849-866); alpha 255
void tonemap_ldr(const float* rgba32f, size_t pixels, float exposure, uint8_t* rgba8);
// minimal PNG writer (8-bit RGBA, stored deflate blocks); returns false on I/O failure
bool write_png(const std::string& path, const uint8_t* rgba8, uint32_t width, uint32_t height);
}  // namespace app
}  // namespace rrh

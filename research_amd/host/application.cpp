// application.cpp -- the headless Application (rr_host.h): pass wiring per display mode,
// deferred scene loads, camera, tonemapped screenshots.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <stdexcept>

#include "rr_host.h"

namespace rrh::app {

bool Application::s_instance_alive_ = false;

Application::Application() = default;
Application::~Application() {
    try {
        shutdown();
    } catch (...) {
    }
}

void Application::initialize() {
    if (initialized_) return;
    // one live (initialized) application per process, like the reference's single window
    if (s_instance_alive_) throw std::runtime_error("Application: only one instance may be initialized at a time");
    device_.create(0);
    s_instance_alive_ = true;
    initialize_renderer();
    initialized_ = true;
    close_requested_ = false;
    load_cornell();  // the default scene, loaded at the first step like every scene request
}

void Application::initialize_renderer() {
    const rhi::Extent2D e{(uint32_t)width_, (uint32_t)height_};
    gbuffer_pass_ = renderer_.add_pass<passes::GBufferPass>();  // first, as Application.cpp:218-221
    pathtracer_pass_ = renderer_.add_pass<passes::PathTracerPass>();
    guiding_pass_ = renderer_.add_pass<passes::GuidingPass>();
    accumulate_pass_ = renderer_.add_pass<passes::AccumulatePass>();
    restir_di_pass_ = renderer_.add_pass<passes::ReSTIRDIPass>();
    restir_gi_pass_ = renderer_.add_pass<passes::ReSTIRGIPass>();  // after DI: its G-buffer and direct image
    restir_pt_pass_ = renderer_.add_pass<passes::ReSTIRPTPass>();
    atrous_pass_ = renderer_.add_pass<passes::AtrousPass>();
    tonemap_pass_ = renderer_.add_pass<passes::TonemapPass>();  // last: the display image
    gbuffer_pass_->initialize(device_, e);
    tonemap_pass_->initialize(device_, e);
    pathtracer_pass_->initialize(device_, e);
    guiding_pass_->initialize(device_, e);
    accumulate_pass_->initialize(device_, e);
    restir_di_pass_->initialize(device_, e);
    restir_gi_pass_->initialize(device_, e);
    restir_gi_pass_->set_inputs(&restir_di_pass_->gbuffer_position(), &restir_di_pass_->gbuffer_normal(),
                                &restir_di_pass_->output_image());
    restir_pt_pass_->initialize(device_, e);
    restir_pt_pass_->set_inputs(&restir_di_pass_->gbuffer_position(), &restir_di_pass_->gbuffer_normal(),
                                &restir_di_pass_->output_image());
    atrous_pass_->initialize(device_, e);
    atrous_pass_->set_input(&restir_di_pass_->output_image());
    atrous_pass_->set_gbuffer(&restir_di_pass_->gbuffer_position(), &restir_di_pass_->gbuffer_normal());
    guiding_pass_->set_enabled(false);
    restir_di_pass_->set_enabled(false);
    restir_gi_pass_->set_enabled(false);
    restir_pt_pass_->set_enabled(false);
    atrous_pass_->set_enabled(false);
}

void Application::shutdown() {
    if (!initialized_) return;
    device_.wait_idle();
    for (auto& p : renderer_.passes()) {
        if (auto* q = dynamic_cast<passes::PathTracerPass*>(p.get())) q->shutdown(device_);
        if (auto* q = dynamic_cast<passes::GuidingPass*>(p.get())) q->shutdown(device_);
        if (auto* q = dynamic_cast<passes::AccumulatePass*>(p.get())) q->shutdown(device_);
        if (auto* q = dynamic_cast<passes::ReSTIRDIPass*>(p.get())) q->shutdown(device_);
        if (auto* q = dynamic_cast<passes::ReSTIRPTPass*>(p.get())) q->shutdown(device_);
        if (auto* q = dynamic_cast<passes::ReSTIRGIPass*>(p.get())) q->shutdown(device_);
        if (auto* q = dynamic_cast<passes::AtrousPass*>(p.get())) q->shutdown(device_);
        if (auto* q = dynamic_cast<passes::GBufferPass*>(p.get())) q->shutdown(device_);
        if (auto* q = dynamic_cast<passes::TonemapPass*>(p.get())) q->shutdown(device_);
    }
    renderer_.clear();
    pathtracer_pass_ = nullptr;
    guiding_pass_ = nullptr;
    accumulate_pass_ = nullptr;
    restir_di_pass_ = nullptr;
    restir_pt_pass_ = nullptr;
    restir_gi_pass_ = nullptr;
    atrous_pass_ = nullptr;
    gbuffer_pass_ = nullptr;
    tonemap_pass_ = nullptr;
    display_ = nullptr;
    scene_.destroy();
    device_.destroy();
    initialized_ = false;
    s_instance_alive_ = false;
}

bool Application::should_close() const {
    return !initialized_ || close_requested_ || (max_frames != 0 && frames_rendered_ >= max_frames);
}

void Application::run_until_closed() {
    if (max_frames == 0 && !close_requested_)
        throw std::runtime_error("run_until_closed: headless application has no window to close; set max_frames");
    while (!should_close()) step();
    device_.wait_idle();
}

void Application::load_cornell() {
    pending_kind_ = RRH_GEN_CORNELL_REF;
    pending_seed_ = 1;
    pending_scale_ = 1.0f;
    pending_gltf_.clear();
}
void Application::load_gltf(const std::string& path) {
    // deferred like the reference's load request (Application.cpp:397-401); empty path ignored
    if (path.empty()) return;
    pending_gltf_ = path;
    pending_kind_ = -1;
}
void Application::load_procedural(int kind, uint32_t seed, float scale) {
    if (kind < RRH_GEN_CORNELL_REF || kind > RRH_GEN_INDOOR) throw std::invalid_argument("load_procedural: bad kind");
    pending_kind_ = kind;
    pending_seed_ = seed;
    pending_scale_ = scale;
    pending_gltf_.clear();
}

void Application::apply_scene_request() {
    if (pending_kind_ < 0 && pending_gltf_.empty()) return;
    device_.wait_idle();
    scene::Camera cam;
    if (!pending_gltf_.empty()) {
        // Application::reload_scene_gltf (Application.cpp:761-793): a file that does not load
        // falls back to the Cornell box; the camera returns to its default
        last_scene_error_.clear();
        try {
            scene::GltfScene g;
            scene::load_gltf(pending_gltf_, g);
            const size_t slash = pending_gltf_.find_last_of('/');
            scene_.upload(device_, g, slash == std::string::npos ? pending_gltf_ : pending_gltf_.substr(slash + 1));
        } catch (const std::exception& e) {
            last_scene_error_ = e.what();
            std::fprintf(stderr, "[Scene] glTF load failed for '%s' (%s), falling back to Cornell Box\n",
                         pending_gltf_.c_str(), e.what());
            scene_.load_generated(device_, RRH_GEN_CORNELL_REF, 1, 1.0f);
        }
        pending_gltf_.clear();
    } else {
        scene_.load_generated(device_, pending_kind_, pending_seed_, pending_scale_);
        cam.fov_y_deg = scene_.suggested_fov;
        cam.look_at(scene_.suggested_eye, scene_.suggested_center);
    }
    current_scene_name_ = scene_.name();
    camera_ = cam;
    pending_kind_ = -1;
    camera_moved_ = true;
    accumulated_spp_ = 0;
    if (restir_di_pass_) restir_di_pass_->reset_history();
    if (restir_pt_pass_) restir_pt_pass_->reset_history();
    if (restir_gi_pass_) restir_gi_pass_->reset_history();
    if (guiding_pass_) guiding_pass_->reset_history();
}

void Application::set_camera_look_at(float ex, float ey, float ez, float cx, float cy, float cz, float ux, float uy,
                                     float uz) {
    camera_.look_at({ex, ey, ez}, {cx, cy, cz}, {ux, uy, uz});
    camera_moved_ = true;  // discard progressive accumulation: the view changed
    accumulated_spp_ = 0;
}
std::array<float, 3> Application::camera_position() const { return camera_.position(); }
std::array<float, 3> Application::camera_forward() const { return camera_.forward(); }

void Application::set_pass_enabled(const std::string& name, bool enabled) {
    if (name == "restir_di") use_di_ = enabled;
    else if (name == "restir_pt") use_pt_ = enabled;
    else if (name == "restir_gi") use_gi_ = enabled;
    else if (name == "denoise") use_denoise_ = enabled;
    else if (name == "guiding") use_guiding_ = enabled;
    else
        throw std::invalid_argument("set_pass_enabled: unknown pass '" + name +
                                    "' (expected restir_di | restir_gi | restir_pt | denoise | guiding)");
}

void Application::set_window_size(int width, int height) {
    if (width <= 0 || height <= 0) throw std::invalid_argument("set_window_size: extent must be positive");
    width_ = width;
    height_ = height;
    if (initialized_) {
        device_.wait_idle();
        renderer_.on_resize({(uint32_t)width, (uint32_t)height});
        camera_moved_ = true;
        accumulated_spp_ = 0;
        if (restir_di_pass_) restir_di_pass_->reset_history();
        if (restir_pt_pass_) restir_pt_pass_->reset_history();
        if (restir_gi_pass_) restir_gi_pass_->reset_history();
    }
}

void Application::step() {
    if (!initialized_) throw std::runtime_error("step: initialize() first");
    apply_scene_request();
    // display modes (Application.cpp:404-441): the real-time ones (DI, ReSTIR PT, optionally
    // denoised) show the frame; path tracing, guided or not, accumulates progressively
    // ReSTIR PT and ReSTIR GI are exclusive indirect modes; PT wins (Application.cpp:420-423)
    const bool use_gi = use_gi_ && !use_pt_;
    const bool use_realtime = use_di_ || use_pt_ || use_gi;
    const bool use_denoise = use_denoise_ && use_realtime;
    if (use_di_ != prev_use_di_ || use_pt_ != prev_use_pt_ || use_gi_ != prev_use_gi_ ||
        use_denoise != prev_use_denoise_ || use_guiding_ != prev_use_guiding_) {
        accumulated_spp_ = 0;
        camera_moved_ = true;
        restir_di_pass_->reset_history();
        restir_pt_pass_->reset_history();
        restir_gi_pass_->reset_history();
        prev_use_di_ = use_di_;
        prev_use_pt_ = use_pt_;
        prev_use_gi_ = use_gi_;
        prev_use_denoise_ = use_denoise;
        prev_use_guiding_ = use_guiding_;
    }
    pathtracer_pass_->set_enabled(!use_realtime && !use_guiding_);
    guiding_pass_->set_enabled(!use_realtime && use_guiding_);
    accumulate_pass_->set_enabled(!use_realtime);
    // ReSTIR PT reads the G-buffer ReSTIR DI writes, so DI runs for it too; its direct image
    // is added only when DI is on (include_direct_lighting = use_di_, Application.cpp:424-425)
    restir_di_pass_->set_enabled(use_di_ || use_pt_ || use_gi);
    restir_pt_pass_->set_enabled(use_pt_);
    restir_pt_pass_->include_direct_lighting = use_di_;
    restir_gi_pass_->set_enabled(use_gi);
    restir_gi_pass_->include_direct_lighting = use_di_;
    atrous_pass_->set_enabled(use_denoise);
    atrous_pass_->set_input(use_pt_ ? &restir_pt_pass_->output_image()
                                    : (use_gi ? &restir_gi_pass_->output_image() : &restir_di_pass_->output_image()));
    accumulate_pass_->set_input(use_guiding_ ? &guiding_pass_->output_image() : &pathtracer_pass_->output_image());
    tonemap_pass_->set_input(current_display_image());
    tonemap_pass_->exposure = exposure_;

    scene_.update_camera(camera_, (uint32_t)width_, (uint32_t)height_, frames_rendered_);
    accumulate_pass_->camera_moved = camera_moved_;
    accumulate_pass_->accumulated_spp = accumulated_spp_;

    render::FrameContext fc;
    fc.device = &device_;
    fc.scene = &scene_;
    fc.extent = {(uint32_t)width_, (uint32_t)height_};
    fc.frame_index = frames_rendered_;
    renderer_.render(fc);
    camera_moved_ = false;
    accumulated_spp_ = accumulate_pass_->accumulated_spp;
    display_ = current_display_image();
    ++frames_rendered_;
}

const rhi::Image* Application::current_display_image() const {
    if (use_pt_) return (use_denoise_ && atrous_pass_) ? &atrous_pass_->output_image() : &restir_pt_pass_->output_image();
    if (use_gi_) return (use_denoise_ && atrous_pass_) ? &atrous_pass_->output_image() : &restir_gi_pass_->output_image();
    if (use_di_) return (use_denoise_ && atrous_pass_) ? &atrous_pass_->output_image() : &restir_di_pass_->output_image();
    return &accumulate_pass_->accumulated_image();
}

std::vector<float> Application::display_image() const {
    if (!display_) throw std::runtime_error("display_image: no frame rendered yet");
    rhi::Device& d = const_cast<rhi::Device&>(device_);
    std::vector<float> hdr((size_t)width_ * height_ * 4);
    display_->download(d, hdr.data(), hdr.size() * sizeof(float));
    return hdr;
}

std::vector<uint8_t> Application::display_ldr() const {
    // the device tonemap at the current exposure (bytes identical to tonemap_ldr on the host)
    if (!display_) throw std::runtime_error("display_ldr: no frame rendered yet");
    rhi::Device& d = const_cast<rhi::Device&>(device_);
    tonemap_pass_->exposure = exposure_;
    tonemap_pass_->run(*display_);
    std::vector<uint8_t> ldr((size_t)width_ * height_ * 4);
    tonemap_pass_->output_image().download(d, ldr.data(), ldr.size());
    return ldr;
}

void Application::capture_screenshot(const std::string& path) {
    device_.wait_idle();
    std::vector<uint8_t> ldr = display_ldr();
    std::string out = path;
    if (out.empty()) out = "screenshot_" + std::to_string(accumulated_spp_) + "spp.png";
    if (!write_png(out, ldr.data(), (uint32_t)width_, (uint32_t)height_))
        throw std::runtime_error("capture_screenshot: cannot write '" + out + "'");
}

// ---- tonemap + PNG ------------------------------------------------------------------------------
void tonemap_ldr(const float* hdr, size_t pixels, float exposure, uint8_t* ldr) {
    // the reference's per-channel formula; the gamma and 8-bit rounding taken exactly through
    // the library's thresholds, so these bytes equal the device tonemap's
    float T[256];
    check(rrh_tonemap_thresholds(T), "rrh_tonemap_thresholds");
    for (size_t i = 0; i < pixels; ++i) {
        for (int c = 0; c < 3; ++c) {
            float v = hdr[i * 4 + c] * exposure;
            v = (v * (2.51f * v + 0.03f)) / (v * (2.43f * v + 0.59f) + 0.14f);
            v = std::fmin(std::fmax(v, 0.0f), 1.0f);
            ldr[i * 4 + c] = (uint8_t)(std::upper_bound(T + 1, T + 256, v) - (T + 1));
        }
        ldr[i * 4 + 3] = 255u;
    }
}

namespace {
uint32_t crc32_update(uint32_t crc, const uint8_t* p, size_t n) {
    static uint32_t table[256];
    static bool init = false;
    if (!init) {  // reflected CRC-32, polynomial 0xEDB88320 (ISO 3309 / PNG specification)
        for (uint32_t k = 0; k < 256; ++k) {
            uint32_t c = k;
            for (int j = 0; j < 8; ++j) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            table[k] = c;
        }
        init = true;
    }
    for (size_t i = 0; i < n; ++i) crc = table[(crc ^ p[i]) & 0xFFu] ^ (crc >> 8);
    return crc;
}
void put32(std::vector<uint8_t>& v, uint32_t x) {
    v.push_back((uint8_t)(x >> 24));
    v.push_back((uint8_t)(x >> 16));
    v.push_back((uint8_t)(x >> 8));
    v.push_back((uint8_t)x);
}
void chunk(std::vector<uint8_t>& out, const char* type, const std::vector<uint8_t>& data) {
    put32(out, (uint32_t)data.size());
    std::vector<uint8_t> td(type, type + 4);
    td.insert(td.end(), data.begin(), data.end());
    out.insert(out.end(), td.begin(), td.end());
    put32(out, crc32_update(0xFFFFFFFFu, td.data(), td.size()) ^ 0xFFFFFFFFu);
}
}  // namespace

bool write_png(const std::string& path, const uint8_t* rgba8, uint32_t width, uint32_t height) {
    // raw scanlines, filter type 0
    std::vector<uint8_t> raw;
    raw.reserve((size_t)height * (width * 4 + 1));
    for (uint32_t y = 0; y < height; ++y) {
        raw.push_back(0);
        raw.insert(raw.end(), rgba8 + (size_t)y * width * 4, rgba8 + (size_t)(y + 1) * width * 4);
    }
    // zlib stream of stored deflate blocks (RFC 1950 / 1951) + Adler-32
    std::vector<uint8_t> z = {0x78, 0x01};
    size_t pos = 0;
    do {
        const size_t n = std::min<size_t>(65535, raw.size() - pos);
        const bool last = pos + n == raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back((uint8_t)(n & 0xFF));
        z.push_back((uint8_t)(n >> 8));
        z.push_back((uint8_t)(~n & 0xFF));
        z.push_back((uint8_t)((~n >> 8) & 0xFF));
        z.insert(z.end(), raw.begin() + pos, raw.begin() + pos + n);
        pos += n;
    } while (pos < raw.size());
    uint32_t a = 1, b = 0;
    for (uint8_t c : raw) {
        a = (a + c) % 65521u;
        b = (b + a) % 65521u;
    }
    put32(z, (b << 16) | a);
    std::vector<uint8_t> png = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    std::vector<uint8_t> ihdr;
    put32(ihdr, width);
    put32(ihdr, height);
    ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});  // 8-bit RGBA, deflate, no filter method, no interlace
    chunk(png, "IHDR", ihdr);
    chunk(png, "IDAT", z);
    chunk(png, "IEND", {});
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(png.data(), 1, png.size(), f) == png.size();
    return (std::fclose(f) == 0) && ok;
}

}  // namespace rrh::app

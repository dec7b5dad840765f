// passes.cpp -- the render passes of the host layer (rr_host.h) over the C ABI.
#include <algorithm>
#include <stdexcept>

#include "rr_host.h"

namespace rrh::passes {

using rhi::Extent2D;
using rhi::Format;
using rhi::Image;

static RenderPass::ResourceDesc desc(const char* name, Format f, Extent2D e, bool persistent) {
    RenderPass::ResourceDesc d;
    d.name = name;
    d.format = f;
    d.extent = e;
    d.persistent = persistent;
    return d;
}

// ---- GBufferPass (GBufferPass.cpp:216-330) ----------------------------------------------------
GBufferPass::~GBufferPass() {
    position_img_.destroy();
    normal_img_.destroy();
    material_id_img_.destroy();
}
void GBufferPass::initialize(rhi::Device& device, Extent2D extent) {
    device_ = &device;
    initialized_ = true;
    on_resize(extent);
}
void GBufferPass::shutdown(rhi::Device&) {
    position_img_.destroy();
    normal_img_.destroy();
    material_id_img_.destroy();
    initialized_ = false;
}
RenderPass::Reflection GBufferPass::reflect() const {
    Reflection r;
    r.outputs.push_back(desc("gbuffer_position", Format::RGBA32F, extent_, false));
    r.outputs.push_back(desc("gbuffer_normal", Format::RGBA32F, extent_, false));
    r.outputs.push_back(desc("gbuffer_material_id", Format::R32U, extent_, false));
    return r;
}
void GBufferPass::on_resize(Extent2D e) {
    if (!initialized_) return;
    extent_ = e;
    position_img_.create(*device_, e, Format::RGBA32F);
    normal_img_.create(*device_, e, Format::RGBA32F);
    material_id_img_.create(*device_, e, Format::R32U);
}
void GBufferPass::execute(FrameContext& fc) {
    if (!initialized_ || !fc.scene || !fc.scene->valid()) return;
    check(rrh_gbuffer(device_->ctx(), fc.scene->handle(), position_img_.handle(), normal_img_.handle(),
                      material_id_img_.handle()),
          "rrh_gbuffer");
}

// ---- TonemapPass (TonemapPass.cpp, tonemap.slang; screenshot Application.cpp:853-866) --------
TonemapPass::~TonemapPass() { ldr_img_.destroy(); }
void TonemapPass::initialize(rhi::Device& device, Extent2D extent) {
    device_ = &device;
    initialized_ = true;
    on_resize(extent);
}
void TonemapPass::shutdown(rhi::Device&) {
    ldr_img_.destroy();
    initialized_ = false;
}
RenderPass::Reflection TonemapPass::reflect() const {
    Reflection r;
    r.outputs.push_back(desc("ldr", Format::RGBA8, extent_, false));
    return r;
}
void TonemapPass::on_resize(Extent2D e) {
    if (!initialized_) return;
    extent_ = e;
    ldr_img_.create(*device_, e, Format::RGBA8);
}
void TonemapPass::run(const Image& hdr) {
    if (!initialized_) throw std::runtime_error("TonemapPass: not initialized");
    check(rrh_tonemap(device_->ctx(), hdr.handle(), exposure, ldr_img_.handle()), "rrh_tonemap");
}
void TonemapPass::execute(FrameContext&) {
    if (!initialized_ || !input_) return;
    run(*input_);
}

// ---- PathTracerPass (PathTracerPass.cpp:170-224) --------------------------------------------
PathTracerPass::~PathTracerPass() { radiance_img_.destroy(); }
void PathTracerPass::initialize(rhi::Device& device, Extent2D extent) {
    device_ = &device;
    extent_ = extent;
    radiance_img_.create(device, extent, Format::RGBA32F);
    initialized_ = true;
}
void PathTracerPass::shutdown(rhi::Device&) {
    radiance_img_.destroy();
    initialized_ = false;
}
RenderPass::Reflection PathTracerPass::reflect() const {
    Reflection r;
    r.outputs.push_back(desc("radiance", Format::RGBA32F, extent_, false));
    return r;
}
void PathTracerPass::on_resize(Extent2D e) {
    if (!initialized_) return;
    extent_ = e;
    radiance_img_.create(*device_, e, Format::RGBA32F);
}
void PathTracerPass::execute(FrameContext& fc) {
    if (!initialized_ || !fc.scene || !fc.scene->valid()) return;  // silent early return, as the reference
    rrh_pt_params p{};
    p.max_bounces = max_bounces;
    p.frame_seed = frame_count_++;
    p.rr_start = rr_start;
    p.guide_mode = RRH_GUIDE_NONE;
    check(rrh_pt_render(device_->ctx(), fc.scene->handle(), &p, radiance_img_.handle()), "rrh_pt_render");
}

// ---- AccumulatePass (AccumulatePass.cpp:167-218) --------------------------------------------
AccumulatePass::~AccumulatePass() {
    accum_img_.destroy();
    accum64_img_.destroy();
}
void AccumulatePass::initialize(rhi::Device& device, Extent2D extent) {
    device_ = &device;
    extent_ = extent;
    accum_img_.create(device, extent, Format::RGBA32F);
    accum64_img_.create(device, extent, Format::RGBA64F);
    initialized_ = true;
}
void AccumulatePass::shutdown(rhi::Device&) {
    accum_img_.destroy();
    accum64_img_.destroy();
    initialized_ = false;
}
RenderPass::Reflection AccumulatePass::reflect() const {
    Reflection r;
    r.inputs.push_back(desc("radiance", Format::RGBA32F, extent_, false));
    r.outputs.push_back(desc("accumulated", Format::RGBA32F, extent_, true));
    if (fp64_enable) r.outputs.push_back(desc("accumulated_sum_f64", Format::RGBA64F, extent_, true));
    return r;
}
void AccumulatePass::on_resize(Extent2D e) {
    if (!initialized_) return;
    extent_ = e;
    accum_img_.create(*device_, e, Format::RGBA32F);
    accum64_img_.create(*device_, e, Format::RGBA64F);
    accumulated_spp = 0;
}
void AccumulatePass::execute(FrameContext&) {
    if (!initialized_ || !input_ || !input_->handle()) return;
    check(rrh_accumulate(device_->ctx(), input_->handle(), accum_img_.handle(), camera_moved ? 1u : 0u,
                         accumulated_spp),
          "rrh_accumulate");
    if (fp64_enable)
        check(rrh_accumulate_f64(device_->ctx(), input_->handle(), accum64_img_.handle(), camera_moved ? 1u : 0u,
                                 accumulated_spp),
              "rrh_accumulate_f64");
    accumulated_spp = camera_moved ? 1u : accumulated_spp + 1u;
    camera_moved = false;
}

// ---- ReSTIRDIPass (ReSTIRDIPass.cpp:306-447) ------------------------------------------------
ReSTIRDIPass::~ReSTIRDIPass() {
    Image* imgs[] = {&gbuf_pos_, &gbuf_norm_, &output_img_, &reservoir_[0], &reservoir_[1], &reservoir_uv_[0],
                     &reservoir_uv_[1]};
    for (Image* i : imgs) i->destroy();
}
void ReSTIRDIPass::initialize(rhi::Device& device, Extent2D extent) {
    device_ = &device;
    extent_ = extent;
    gbuf_pos_.create(device, extent, Format::RGBA32F);
    gbuf_norm_.create(device, extent, Format::RGBA32F);
    output_img_.create(device, extent, Format::RGBA32F);
    for (int k = 0; k < 2; ++k) {
        reservoir_[k].create(device, extent, Format::RGBA32F);
        reservoir_uv_[k].create(device, extent, Format::R32U);
    }
    reset_history();
    initialized_ = true;
}
void ReSTIRDIPass::shutdown(rhi::Device&) {
    Image* imgs[] = {&gbuf_pos_, &gbuf_norm_, &output_img_, &reservoir_[0], &reservoir_[1], &reservoir_uv_[0],
                     &reservoir_uv_[1]};
    for (Image* i : imgs) i->destroy();
    initialized_ = false;
}
RenderPass::Reflection ReSTIRDIPass::reflect() const {
    Reflection r;
    r.outputs.push_back(desc("gbuffer_position", Format::RGBA32F, extent_, false));
    r.outputs.push_back(desc("gbuffer_normal", Format::RGBA32F, extent_, false));
    r.outputs.push_back(desc("reservoir_a", Format::RGBA32F, extent_, true));
    r.outputs.push_back(desc("reservoir_b", Format::RGBA32F, extent_, true));
    r.outputs.push_back(desc("direct_radiance", Format::RGBA32F, extent_, false));
    return r;
}
void ReSTIRDIPass::on_resize(Extent2D e) {
    if (!initialized_) return;
    rhi::Device& d = *device_;
    shutdown(d);
    initialize(d, e);
}
void ReSTIRDIPass::execute(FrameContext& fc) {
    if (!initialized_ || !fc.scene || !fc.scene->valid()) return;
    rrh_ctx* c = device_->ctx();
    rrh_scene* s = fc.scene->handle();
    const uint32_t cur = reservoir_flip_, prev = 1u - reservoir_flip_;
    rrh_di_params p{};
    p.frame_seed = frame_count_++;
    p.num_initial_candidates = num_initial_candidates;
    p.temporal_enable = temporal_enable ? 1u : 0u;
    p.spatial_enable = spatial_enable ? 1u : 0u;
    p.spatial_num_neighbors = spatial_num_neighbors;
    p.spatial_radius = spatial_radius;
    if (first_execute_) {  // no history yet: the previous reservoirs start empty
        reservoir_[prev].clear(*device_);
        reservoir_uv_[prev].clear(*device_);
    }
    check(rrh_di_initial(c, s, &p, gbuf_pos_.handle(), gbuf_norm_.handle(), reservoir_[cur].handle(),
                         reservoir_uv_[cur].handle()),
          "rrh_di_initial");
    check(rrh_di_temporal(c, s, &p, gbuf_pos_.handle(), gbuf_norm_.handle(), reservoir_[cur].handle(),
                          reservoir_uv_[cur].handle(), reservoir_[prev].handle(), reservoir_uv_[prev].handle()),
          "rrh_di_temporal");
    check(rrh_di_spatial_shade(c, s, &p, gbuf_pos_.handle(), gbuf_norm_.handle(), reservoir_[cur].handle(),
                               reservoir_uv_[cur].handle(), output_img_.handle()),
          "rrh_di_spatial_shade");
    reservoir_flip_ = prev;
    first_execute_ = false;
}

// ---- ReSTIRPTPass (ReSTIRPTPass.cpp:529-813) ------------------------------------------------
ReSTIRPTPass::~ReSTIRPTPass() {
    destroy_images();
    if (grid_) rrh_grid_destroy(grid_);
    grid_ = nullptr;
}
void ReSTIRPTPass::destroy_images() {
    for (auto& r : reservoir_)
        for (Image& i : r) i.destroy();
    for (int k = 0; k < 2; ++k) {
        history_pos_[k].destroy();
        history_norm_[k].destroy();
    }
    output_img_.destroy();
}
void ReSTIRPTPass::initialize(rhi::Device& device, Extent2D extent) {
    device_ = &device;
    extent_ = extent;
    for (auto& r : reservoir_)
        for (Image& i : r) i.create(device, extent, Format::RGBA32F);
    for (int k = 0; k < 2; ++k) {
        history_pos_[k].create(device, extent, Format::RGBA32F);
        history_norm_[k].create(device, extent, Format::RGBA32F);
    }
    output_img_.create(device, extent, Format::RGBA32F);
    reset_history();
    initialized_ = true;
}
void ReSTIRPTPass::shutdown(rhi::Device&) {
    destroy_images();
    if (grid_) rrh_grid_destroy(grid_);
    grid_ = nullptr;
    initialized_ = false;
}
RenderPass::Reflection ReSTIRPTPass::reflect() const {
    Reflection r;
    r.inputs.push_back(desc("gbuffer_position", Format::RGBA32F, extent_, false));
    r.inputs.push_back(desc("gbuffer_normal", Format::RGBA32F, extent_, false));
    r.inputs.push_back(desc("direct_radiance", Format::RGBA32F, extent_, false));
    for (const char* n : {"reservoir_pos", "reservoir_nrm", "reservoir_rad", "reservoir_pre", "history_primary_pos",
                          "history_primary_norm"})
        r.outputs.push_back(desc(n, Format::RGBA32F, extent_, true));
    r.outputs.push_back(desc("output", Format::RGBA32F, extent_, false));
    return r;
}
void ReSTIRPTPass::on_resize(Extent2D e) {
    if (!initialized_) return;
    destroy_images();
    initialize(*device_, e);
}
void ReSTIRPTPass::execute(FrameContext& fc) {
    if (!initialized_ || !fc.scene || !fc.scene->valid() || !gbuf_pos_ || !gbuf_norm_) return;
    rrh_ctx* c = device_->ctx();
    rrh_scene* s = fc.scene->handle();
    // the guiding grid (ReSTIRPTPass.cpp:311-341): made on first use, made again when its
    // knobs change; a reset clears the accumulator and every fitted lobe (:602-612)
    if (guiding_enable && (!grid_ || grid_made_[0] != grid_base_cell_size || grid_made_[1] != grid_distance_slope ||
                           grid_made_[2] != grid_kappa_max || grid_made_min_ != grid_min_samples)) {
        if (grid_) rrh_grid_destroy(grid_);
        grid_ = nullptr;
        rrh_grid_params g{};
        g.kind = RRH_GRID_HASH;
        g.cell_count = kGridCellCount;
        g.base_cell_size = grid_base_cell_size;
        g.distance_slope = grid_distance_slope;
        g.min_samples = grid_min_samples;
        g.kappa_max = grid_kappa_max;
        check(rrh_grid_create(c, &g, &grid_), "rrh_grid_create");
        grid_made_[0] = grid_base_cell_size;
        grid_made_[1] = grid_distance_slope;
        grid_made_[2] = grid_kappa_max;
        grid_made_min_ = grid_min_samples;
        grid_dirty_ = true;
    }
    if (grid_dirty_) {
        if (grid_) check(rrh_grid_reset(c, grid_), "rrh_grid_reset");
        grid_dirty_ = false;
    }
    const uint32_t cur = reservoir_flip_, prev = 1u - reservoir_flip_;
    rrh_image* res_cur[4];
    rrh_image* res_prev[4];
    for (int k = 0; k < 4; ++k) {
        res_cur[k] = reservoir_[cur][k].handle();
        res_prev[k] = reservoir_[prev][k].handle();
    }
    if (history_dirty_) {  // an explicitly invalid history (ReSTIRPTPass.cpp:616-663)
        check(rrh_rpt_clear(c, res_prev, history_pos_[prev].handle(), history_norm_[prev].handle()), "rrh_rpt_clear");
        history_dirty_ = false;
    }
    rrh_rpt_params p{};
    p.frame_seed = frame_count_++;
    p.num_initial_candidates = num_initial_candidates;
    p.max_bounces = max_bounces;
    p.temporal_enable = temporal_enable ? 1u : 0u;
    p.spatial_enable = spatial_enable ? 1u : 0u;
    p.spatial_num_neighbors = spatial_num_neighbors;
    p.spatial_radius = spatial_radius;
    p.include_direct_lighting = include_direct_lighting ? 1u : 0u;
    p.indirect_strength = indirect_strength;
    p.ray_bias = ray_bias;
    p.guiding_enable = guiding_enable ? 1u : 0u;
    p.guiding_alpha = guiding_alpha;
    rrh_grid* g = guiding_enable ? grid_ : nullptr;
    check(rrh_rpt_initial(c, s, &p, g, gbuf_pos_->handle(), gbuf_norm_->handle(), res_cur, history_pos_[cur].handle(),
                          history_norm_[cur].handle()),
          "rrh_rpt_initial");
    check(rrh_rpt_temporal(c, s, &p, gbuf_pos_->handle(), gbuf_norm_->handle(), res_cur, res_prev,
                           history_pos_[prev].handle(), history_norm_[prev].handle()),
          "rrh_rpt_temporal");
    check(rrh_rpt_spatial_shade(c, s, &p, g, gbuf_pos_->handle(), gbuf_norm_->handle(),
                                (include_direct_lighting && direct_) ? direct_->handle() : nullptr, res_cur,
                                output_img_.handle()),
          "rrh_rpt_spatial_shade");
    if (g) check(rrh_grid_fit(c, g), "rrh_grid_fit");  // lobes for the next frame (:784-810)
    reservoir_flip_ ^= 1u;
}

// ---- ReSTIRGIPass (ReSTIRGIPass.cpp:428-651) ------------------------------------------------
ReSTIRGIPass::~ReSTIRGIPass() { destroy_images(); }
void ReSTIRGIPass::destroy_images() {
    for (auto& r : reservoir_)
        for (Image& i : r) i.destroy();
    for (int k = 0; k < 2; ++k) {
        history_pos_[k].destroy();
        history_norm_[k].destroy();
    }
    output_img_.destroy();
}
void ReSTIRGIPass::initialize(rhi::Device& device, Extent2D extent) {
    device_ = &device;
    extent_ = extent;
    for (auto& r : reservoir_)
        for (Image& i : r) i.create(device, extent, Format::RGBA32F);
    for (int k = 0; k < 2; ++k) {
        history_pos_[k].create(device, extent, Format::RGBA32F);
        history_norm_[k].create(device, extent, Format::RGBA32F);
    }
    output_img_.create(device, extent, Format::RGBA32F);
    reset_history();
    initialized_ = true;
}
void ReSTIRGIPass::shutdown(rhi::Device&) {
    destroy_images();
    initialized_ = false;
}
RenderPass::Reflection ReSTIRGIPass::reflect() const {
    Reflection r;
    r.inputs.push_back(desc("gbuffer_position", Format::RGBA32F, extent_, false));
    r.inputs.push_back(desc("gbuffer_normal", Format::RGBA32F, extent_, false));
    r.inputs.push_back(desc("direct_radiance", Format::RGBA32F, extent_, false));
    for (const char* n : {"reservoir_pos", "reservoir_rad", "reservoir_meta", "history_primary_pos",
                          "history_primary_norm"})
        r.outputs.push_back(desc(n, Format::RGBA32F, extent_, true));
    r.outputs.push_back(desc("output", Format::RGBA32F, extent_, false));
    return r;
}
void ReSTIRGIPass::on_resize(Extent2D e) {
    if (!initialized_) return;
    destroy_images();
    initialize(*device_, e);
}
void ReSTIRGIPass::execute(FrameContext& fc) {
    if (!initialized_ || !fc.scene || !fc.scene->valid() || !gbuf_pos_ || !gbuf_norm_) return;
    rrh_ctx* c = device_->ctx();
    rrh_scene* s = fc.scene->handle();
    const uint32_t cur = reservoir_flip_, prev = 1u - reservoir_flip_;
    rrh_image* res_cur[3];
    rrh_image* res_prev[3];
    for (int k = 0; k < 3; ++k) {
        res_cur[k] = reservoir_[cur][k].handle();
        res_prev[k] = reservoir_[prev][k].handle();
    }
    if (history_dirty_) {
        check(rrh_rgi_clear(c, res_prev, history_pos_[prev].handle(), history_norm_[prev].handle()), "rrh_rgi_clear");
        history_dirty_ = false;
    }
    rrh_rgi_params p{};
    p.frame_seed = frame_count_++;
    p.num_initial_candidates = num_initial_candidates;
    p.temporal_enable = temporal_enable ? 1u : 0u;
    p.spatial_enable = spatial_enable ? 1u : 0u;
    p.spatial_num_neighbors = spatial_num_neighbors;
    p.spatial_radius = spatial_radius;
    p.include_direct_lighting = include_direct_lighting ? 1u : 0u;
    p.indirect_strength = indirect_strength;
    p.ray_bias = ray_bias;
    check(rrh_rgi_initial(c, s, &p, gbuf_pos_->handle(), gbuf_norm_->handle(), res_cur, history_pos_[cur].handle(),
                          history_norm_[cur].handle()),
          "rrh_rgi_initial");
    check(rrh_rgi_temporal(c, s, &p, gbuf_pos_->handle(), gbuf_norm_->handle(), res_cur, res_prev,
                           history_pos_[prev].handle(), history_norm_[prev].handle()),
          "rrh_rgi_temporal");
    check(rrh_rgi_spatial_shade(c, s, &p, gbuf_pos_->handle(), gbuf_norm_->handle(),
                                (include_direct_lighting && direct_) ? direct_->handle() : nullptr, res_cur,
                                output_img_.handle()),
          "rrh_rgi_spatial_shade");
    reservoir_flip_ ^= 1u;
}

// ---- AtrousPass (AtrousPass.cpp:243-306) ----------------------------------------------------
AtrousPass::~AtrousPass() {
    ping_images_[0].destroy();
    ping_images_[1].destroy();
}
void AtrousPass::initialize(rhi::Device& device, Extent2D extent) {
    device_ = &device;
    extent_ = extent;
    ping_images_[0].create(device, extent, Format::RGBA32F);
    ping_images_[1].create(device, extent, Format::RGBA32F);
    out_slot_ = -1;
    initialized_ = true;
}
void AtrousPass::shutdown(rhi::Device&) {
    ping_images_[0].destroy();
    ping_images_[1].destroy();
    initialized_ = false;
}
RenderPass::Reflection AtrousPass::reflect() const {
    Reflection r;
    r.inputs.push_back(desc("radiance", Format::RGBA32F, extent_, false));
    r.inputs.push_back(desc("gbuffer_position", Format::RGBA32F, extent_, false));
    r.inputs.push_back(desc("gbuffer_normal", Format::RGBA32F, extent_, false));
    r.outputs.push_back(desc("ping", Format::RGBA32F, extent_, false));
    r.outputs.push_back(desc("pong", Format::RGBA32F, extent_, false));
    return r;
}
void AtrousPass::on_resize(Extent2D e) {
    if (!initialized_) return;
    rhi::Device& d = *device_;
    shutdown(d);
    initialize(d, e);
}
void AtrousPass::execute(FrameContext&) {
    if (!initialized_ || !input_ || !gbuf_pos_ || !gbuf_norm_) return;
    rrh_atrous_params p{};
    p.iterations = iterations;
    p.mode = RRH_ATROUS_REFERENCE;
    p.sigma_position = sigma_position;
    p.sigma_normal = sigma_normal;
    p.sigma_luminance = sigma_luminance;
    p.flags = fast_math ? RRH_ATROUS_FAST_MATH : 0u;
    int slot = 0;
    check(rrh_atrous(device_->ctx(), &p, input_->handle(), gbuf_pos_->handle(), gbuf_norm_->handle(),
                     ping_images_[0].handle(), ping_images_[1].handle(), &slot),
          "rrh_atrous");
    out_slot_ = slot;
}

// ---- GuidingPass ------------------------------------------------------------------------------
GuidingPass::~GuidingPass() {
    destroy_guide();
    radiance_img_.destroy();
}
void GuidingPass::initialize(rhi::Device& device, Extent2D extent) {
    device_ = &device;
    extent_ = extent;
    radiance_img_.create(device, extent, Format::RGBA32F);
    initialized_ = true;
}
void GuidingPass::shutdown(rhi::Device&) {
    destroy_guide();
    radiance_img_.destroy();
    initialized_ = false;
}
RenderPass::Reflection GuidingPass::reflect() const {
    Reflection r;
    r.outputs.push_back(desc("radiance", Format::RGBA32F, extent_, false));
    return r;
}
void GuidingPass::on_resize(Extent2D e) {
    if (!initialized_) return;
    extent_ = e;
    radiance_img_.create(*device_, e, Format::RGBA32F);
}
void GuidingPass::destroy_guide() {
    if (grid_) rrh_grid_destroy(grid_);
    if (mlp_) rrh_mlp_destroy(mlp_);
    grid_ = nullptr;
    mlp_ = nullptr;
    guide_scene_ = nullptr;
}
void GuidingPass::reset_history() {
    destroy_guide();
    frame_count_ = 0;
}
void GuidingPass::create_guide(scene::Scene& s) {
    destroy_guide();
    const rrh_scene_info info = s.info();
    if (mode == Mode::VMM) {
        rrh_grid_params g{};
        g.kind = RRH_GRID_VMM;
        g.kappa_max = grid_kappa_max;
        g.res = grid_resolution;
        g.lobes = grid_lobes;
        for (int k = 0; k < 3; ++k) {
            g.bounds_min[k] = info.bounds_min[k];
            g.bounds_max[k] = info.bounds_max[k];
        }
        g.decay = grid_decay;
        check(rrh_grid_create(device_->ctx(), &g, &grid_), "rrh_grid_create");
    } else {
        rrh_mlp_params m{};
        for (int k = 0; k < 3; ++k) {
            m.bounds_min[k] = info.bounds_min[k];
            m.bounds_max[k] = info.bounds_max[k];
        }
        m.lr = mlp_learning_rate;
        m.beta1 = 0.9f;
        m.beta2 = 0.999f;
        m.eps = 1e-8f;
        m.seed = 1;
        m.loss_scale = 1024.0f;
        check(rrh_mlp_create(device_->ctx(), &m, &mlp_), "rrh_mlp_create");
    }
    guide_scene_ = s.handle();
    guide_mode_ = mode;
    frame_count_ = 0;
}
void GuidingPass::execute(FrameContext& fc) {
    if (!initialized_ || !fc.scene || !fc.scene->valid()) return;
    if (guide_scene_ != fc.scene->handle() || guide_mode_ != mode || (!grid_ && !mlp_)) create_guide(*fc.scene);
    rrh_ctx* c = device_->ctx();
    rrh_pt_params p{};
    p.max_bounces = max_bounces;
    p.frame_seed = frame_count_;
    p.rr_start = 3;
    p.guide_mode = guiding_enable ? (mode == Mode::VMM ? RRH_GUIDE_VMM : RRH_GUIDE_MLP) : RRH_GUIDE_NONE;
    p.guide_alpha = (frame_count_ > 0) ? guiding_alpha : 0.0f;  // nothing learned before the first frame
    p.record_vertices = guiding_enable ? (mode == Mode::MLP ? mlp_batch : record_vertices) : 0u;
    check(rrh_pt_render_guided(c, fc.scene->handle(), &p, grid_, mlp_, radiance_img_.handle()), "rrh_pt_render_guided");
    if (guiding_enable) {
        if (mode == Mode::VMM) {
            check(rrh_grid_em_from_vertices(c, grid_), "rrh_grid_em_from_vertices");
        } else {
            for (uint32_t k = 0; k < mlp_steps_per_frame; ++k)
                check(rrh_mlp_train_from_vertices(c, mlp_, mlp_batch, k + 1 == mlp_steps_per_frame ? &last_loss_ : nullptr),
                      "rrh_mlp_train_from_vertices");
        }
    }
    ++frame_count_;
}

}  // namespace rrh::passes

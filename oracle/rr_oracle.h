/*
 * rr_oracle.h -- CPU restatement of the reference renderer's hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it may be loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * by nothing else.  The product (research_amd/, librrh.so) never links it.
 *
 * Every function restates one piece of the reference's Slang/C++ code and
 * cites the file:line it follows (paths relative to the reference root).
 * Arithmetic is IEEE fp32 evaluated in source order (built with
 * -ffp-contract=off, no -ffast-math) so the GPU kernels, which follow the
 * same operation order, can be compared with it bit for bit.
 *
 * Transcendentals (sin, cos, exp, log) are the project's own deterministic
 * implementations (rr_sinf etc. below): Cody-Waite range reduction plus
 * Taylor / atanh series.  The reference calls the Vulkan driver's
 * transcendentals, whose ULP behaviour is driver-specific and cannot be
 * pinned; using the same explicit routines on CPU and GPU makes the oracle
 * and the kernels comparable exactly.
 */
#ifndef RR_ORACLE_H
#define RR_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- scene records: byte layouts of src/scene/GpuScene.h:17-114 ---------- */
typedef struct { float pos[3]; float normal[3]; float uv[2]; float tangent[4]; } orc_vertex;      /* 48 B */
typedef struct { uint32_t vertex_base, vertex_count, index_byte_offset, index_count,
                 material_index, _pad[3]; } orc_mesh;                                              /* 32 B */
typedef struct { float base_color[4]; float metallic, roughness; float emissive[3]; float ior;
                 uint32_t albedo_tex_idx, normal_tex_idx, metallic_roughness_tex_idx,
                          emissive_tex_idx, sampler_idx; uint32_t flags, _pad1, _pad2; } orc_material; /* 72 B */
typedef struct { float transform[16]; uint32_t mesh_index; uint32_t _pad[3]; } orc_instance;       /* 80 B */
typedef struct { float position[3]; uint32_t type; float direction[3]; float radius;
                 float emission[3]; float intensity; } orc_light;                                  /* 48 B */
typedef struct { float view[16], inv_view[16], inv_proj[16], prev_view_proj[16];
                 float position[3]; float near_plane; float direction[3]; float far_plane;
                 uint32_t frame_index, screen_width, screen_height, _pad; } orc_camera;            /* 304 B */

/* material flags (the reference's GpuMaterial._pad[0]) */
#define ORC_MAT_PURE_LAMBERT 1u

/* scene flags */
#define ORC_SCENE_TRI_LIGHTS 1u   /* emissive triangles are NEE / ReSTIR lights (MIS on BSDF hits) */

typedef struct orc_scene orc_scene;

/* ---- deterministic math ---------------------------------------------------- */
float orc_sinf(float x);
float orc_cosf(float x);
float orc_expf(float x);
float orc_wexpf(float x);  /* the A-trous config mode's weight exp */
float orc_logf(float x);

/* ---- RNG: Common.slang:18-28, Sampling.slang:11-30 ------------------------- */
uint32_t orc_pcg_hash(uint32_t seed);
float    orc_uint_to_float01(uint32_t v);
float    orc_rand_float(uint32_t* seed);
uint32_t orc_make_seed(uint32_t px, uint32_t py, uint32_t frame);

/* ---- camera: Camera.cpp:27-146 + Scene.cpp:554-585 (restated in double) ---- */
void orc_camera_look_at(const float eye[3], const float center[3], float fov_y_deg,
                        uint32_t width, uint32_t height, float near_plane, float far_plane,
                        uint32_t frame_index, orc_camera* out);
void orc_primary_ray(const orc_camera* cam, uint32_t px, uint32_t py, uint32_t* seed,
                     float org[3], float dir[3]);

/* ---- scene ------------------------------------------------------------------- */
orc_scene* orc_scene_create(const orc_vertex* v, uint32_t nv, const uint32_t* idx, uint32_t ni,
                            const orc_mesh* m, uint32_t nm, const orc_material* mat, uint32_t nmat,
                            const orc_instance* inst, uint32_t ninst, const orc_light* l, uint32_t nl,
                            uint32_t flags);
void     orc_scene_destroy(orc_scene* s);
uint32_t orc_scene_tri_count(const orc_scene* s);
uint32_t orc_scene_light_count(const orc_scene* s);  /* reference lights + triangle lights */
/* world-space triangles, 9 floats each, in flattening order (instance-major) */
void     orc_scene_triangles(const orc_scene* s, float* out);

/* ---- intersection ------------------------------------------------------------ */
typedef struct { float t; uint32_t prim; float b1, b2; } orc_hit;   /* prim = ~0u on miss */
/* watertight (Woop, Benthin, Wald 2013) two-sided closest hit; brute force over all tris */
void orc_trace_brute(const orc_scene* s, const float* rays /* n x 8: o3,tmin,d3,tmax */,
                     uint32_t n, orc_hit* hits);
/* same result through the oracle's own median-split BVH */
void orc_trace_bvh(const orc_scene* s, const float* rays, uint32_t n, orc_hit* hits);
/* G-buffer by pixel-centre rays, front faces only (GBufferPass.cpp:216-330, gbuffer.slang) */
void orc_gbuffer(const orc_scene* s, const orc_camera* cam, float* pos, float* nrm, uint32_t* mat);
/* ACES + gamma 8-bit screenshot values (Application.cpp:853-866), with powf as written */
void orc_tonemap(const float* hdr, size_t pixels, float exposure, uint8_t* ldr);
int  orc_occluded(const orc_scene* s, const float o[3], const float d[3], float tmin, float tmax);

/* ---- BRDF: BRDF.slang:16-187 ------------------------------------------------- */
void  orc_brdf_eval(const float n[3], const float wo[3], const float wi[3],
                    const float base[3], float metallic, float roughness, float out[3]);
void  orc_brdf_sample(const float wo[3], const float n[3], const float base[3], float metallic,
                      float roughness, float u0, float u1, float wi[3], float* pdf, float f[3]);
float orc_brdf_pdf(const float wo[3], const float wi[3], const float n[3], float roughness);
void  orc_sample_cosine_hemisphere(float u0, float u1, float out[3], float* pdf);
void  orc_build_onb(const float n[3], float T[3], float B[3]);
/* BRDF.slang pieces, and array forms for quadrature / histogram checks */
void  orc_fresnel_schlick(float cos_theta, const float F0[3], float out[3]);
float orc_distribution_ggx(float NdotH, float alpha);
float orc_geometry_smith_ggx(float NdotV, float NdotL, float alpha);
void  orc_sample_ggx_vndf_n(const float wo_local[3], float roughness, const float* u, uint32_t count, float* h);
void  orc_brdf_eval_n(const float n[3], const float wo[3], const float* wi, uint32_t count, const float base[3],
                      float metallic, float roughness, float* out);
void  orc_brdf_pdf_n(const float n[3], const float wo[3], const float* wi, uint32_t count, float roughness,
                     float* out);
void  orc_brdf_sample_n(const float n[3], const float wo[3], const float base[3], float metallic, float roughness,
                        const float* u, uint32_t count, float* wi, float* pdf, float* f);

/* ---- path tracer: pathtracer.slang:181-265 ----------------------------------- */
typedef struct {
    uint32_t max_bounces;
    uint32_t frame_seed;
    uint32_t rr_start;       /* bounce index from which Russian roulette runs (3) */
    uint32_t guide_mode;     /* 0 none, 1 vMF grid (8-lobe) */
    float    guide_alpha;
    uint32_t _pad[3];
} orc_pt_params;
typedef struct { uint64_t primary, extension, shadow; } orc_ray_counts;
/* renders the listed pixels (pixel ids y*W+x) into rgba (4 floats per listed pixel) */
void orc_pt_render(const orc_scene* s, const orc_camera* cam, const orc_pt_params* p,
                   const uint32_t* pixels, uint32_t npix, float* rgba, orc_ray_counts* counts,
                   int threads);

/* ---- accumulate: accumulate.slang:21-46 -------------------------------------- */
void orc_accumulate(const float* radiance, float* accum, size_t npix, uint32_t camera_moved,
                    uint32_t accumulated_spp);
void orc_accumulate_f64(const float* radiance, double* accum, size_t npix, uint32_t camera_moved,
                        uint32_t accumulated_spp);

/* ---- ReSTIR DI: restir_di.slang:57-472 ---------------------------------------- */
typedef struct {
    uint32_t frame_seed, num_initial_candidates, temporal_enable, spatial_enable;
    uint32_t spatial_num_neighbors, spatial_radius, _pad0, _pad1;
} orc_di_params;
/* images are W*H float4 (gbuf_pos, gbuf_norm, reservoir) + W*H uint32 (reservoir light uv) */
void orc_di_initial(const orc_scene* s, const orc_camera* cam, const orc_di_params* p,
                    float* gbuf_pos, float* gbuf_norm, float* res, uint32_t* res_uv, int threads);
void orc_di_temporal(const orc_scene* s, const orc_camera* cam, const orc_di_params* p,
                     const float* gbuf_pos, const float* gbuf_norm, float* res_curr,
                     uint32_t* res_curr_uv, const float* res_prev, const uint32_t* res_prev_uv,
                     int threads);
void orc_di_spatial_shade(const orc_scene* s, const orc_camera* cam, const orc_di_params* p,
                          const float* gbuf_pos, const float* gbuf_norm, const float* res,
                          const uint32_t* res_uv, float* out, int threads);
/* shadow rays traced by the last orc_di_spatial_shade call (one per shaded pixel whose sample faces it) */
uint64_t orc_di_last_shadow_rays(void);

/* ---- ReSTIR PT, reconnection shift: restir_pt.slang:98-1001 -------------------- */
/* ReSTIR PT and GI rays are two-sided by default, as the reference's CULL_DISABLE instances
 * trace them; 1 culls back faces as RAY_FLAG_CULL_BACK_FACING_TRIANGLES alone would */
void orc_set_restir_cull(int on);
typedef struct {
    uint32_t frame_seed, num_initial_candidates, max_bounces, temporal_enable, spatial_enable;
    uint32_t spatial_num_neighbors, spatial_radius, include_direct_lighting;
    float indirect_strength, ray_bias;
    uint32_t guiding_enable;
    float guiding_alpha;
} orc_rpt_params;
/* res: four planes of W*H float4 back to back (reconnection position + w_sum, normal + M,
 * suffix + W, prefix + valid); fitted: the hash grid's lobes (4 floats per cell) or NULL;
 * accum: its splat accumulator (4 uint32 per cell) or NULL */
void orc_rpt_initial(const orc_scene* s, const orc_camera* cam, const orc_rpt_params* p, const float* fitted,
                     float base_cell, float slope, uint32_t cell_count, const float* gbuf_pos,
                     const float* gbuf_norm, float* res, float* hist_pos, float* hist_norm, int threads);
void orc_rpt_temporal(const orc_scene* s, const orc_camera* cam, const orc_rpt_params* p, const float* gbuf_pos,
                      const float* gbuf_norm, float* res, const float* res_prev, const float* hist_prev_pos,
                      const float* hist_prev_norm, int threads);
void orc_rpt_spatial_shade(const orc_scene* s, const orc_camera* cam, const orc_rpt_params* p, uint32_t* accum,
                           float base_cell, float slope, uint32_t cell_count, const float* gbuf_pos,
                           const float* gbuf_norm, const float* direct, const float* res, float* out, int threads);

/* ---- ReSTIR GI: restir_gi.slang:143-1013 --------------------------------------- */
typedef struct {
    uint32_t frame_seed, num_initial_candidates, temporal_enable, spatial_enable;
    uint32_t spatial_num_neighbors, spatial_radius, include_direct_lighting;
    float indirect_strength, ray_bias;
    uint32_t _pad[3];
} orc_rgi_params;
/* res: three planes of W*H float4 back to back (sample position + w_sum, radiance + W,
 * normal + M) */
void orc_rgi_initial(const orc_scene* s, const orc_camera* cam, const orc_rgi_params* p, const float* gbuf_pos,
                     const float* gbuf_norm, float* res, float* hist_pos, float* hist_norm, int threads);
void orc_rgi_temporal(const orc_scene* s, const orc_camera* cam, const orc_rgi_params* p, const float* gbuf_pos,
                      const float* gbuf_norm, float* res, const float* res_prev, const float* hist_prev_pos,
                      const float* hist_prev_norm, int threads);
void orc_rgi_spatial_shade(const orc_scene* s, const orc_camera* cam, const orc_rgi_params* p, const float* gbuf_pos,
                           const float* gbuf_norm, const float* direct, const float* res, float* out, int threads);

/* ---- A-trous: atrous.slang:21-106 (reference mode) and SVGF-style config mode ---- */
typedef struct {
    uint32_t width, height, stride, mode;       /* mode 0 reference, 1 config (fp16 I/O) */
    float sigma_position, sigma_normal, sigma_luminance, _pad;
} orc_atrous_params;
/* reference mode: in/out RGBA32F, gpos RGBA32F (w<0 = miss), gnrm RGBA32F */
void orc_atrous_ref(const orc_atrous_params* p, const float* in, const float* gpos,
                    const float* gnrm, float* out);
/* config mode: in/out RGBA16F (rgb + luminance variance in a), gbuf RGBA16F (n.xyz, linear depth; depth<=0 = miss) */
void orc_atrous_cfg(const orc_atrous_params* p, const uint16_t* in, const uint16_t* gbuf,
                    uint16_t* out);
/* config-mode inputs (fp16 radiance + 3x3 luminance variance, fp16 normal + linear depth) from a
 * rendered frame, as rrh_atrous_prepare */
void orc_atrous_prepare(uint32_t W, uint32_t H, const float cpos[3], const float cdir[3], const float* rad,
                        const float* gpos, const float* gnrm, uint16_t* rad16, uint16_t* gb16);
uint16_t orc_f32_to_f16(float f);
float    orc_f16_to_f32(uint16_t h);

/* ---- vMF guiding grid: GuidingGrid.slang:54-189, restir_pt.slang:520-563 ----- */
uint32_t orc_guiding_cell_index(const float p[3], uint32_t bsdf_id, const float cam[3],
                                float base_cell, float slope, uint32_t cell_count);
void     orc_guiding_splat(uint32_t* accum, uint32_t cell, const float dir[3]);
void     orc_guiding_fit_cell(uint32_t* accum, float* fitted, uint32_t cell,
                              uint32_t min_samples, float kappa_max);
void     orc_vmf_sample(const float mu[3], float kappa, float u0, float u1, float dir[3], float* pdf);
float    orc_vmf_pdf(const float mu[3], float kappa, const float dir[3]);
void     orc_sample_primary_dir(const float* fitted, const float pos[3], const float n[3],
                                uint32_t bsdf_id, const float cam[3], float base_cell, float slope,
                                uint32_t cell_count, uint32_t guiding_enable, float alpha,
                                uint32_t* seed, float wi[3], float* src_pdf);

/* ---- 8-lobe vMF mixture EM (new; K=1 M-step reduces to the reference fit) ---- */
/* lobes: per cell K x (mu3, kappa, pi) ; stats: per cell K x (s0, s1x, s1y, s1z) + per cell (wsum) */
void orc_vmm_estep(const float* lobes, uint32_t K, const float* pos, const float* dir,
                   const float* w, uint32_t n, const float bmin[3], const float bmax[3],
                   uint32_t res, double* stats);
void orc_vmm_mstep(const double* stats, uint32_t ncell, uint32_t K, float kappa_max,
                   double min_weight, float* lobes);
double orc_vmm_pdf(const float* lobes_cell, uint32_t K, const float dir[3]);

/* ---- learned guide (FP64; new): see rr_oracle_mlp.c for the definition ---- */
uint32_t orc_mlp_param_count(void);
void   orc_mlp_encode(const float pos[3], const float bmin[3], const float bmax[3], const float n[3],
                      const float wo[3], double* x78);
void   orc_mlp_forward(const double* params, const double* x78, double* acts256, double* out40);
void   orc_mlp_head(const double* out40, double* lobes40);
double orc_mlp_pdf(const double* lobes40, const float dir[3]);
double orc_mlp_loss_grad(const double* params, const double* x, const float* dirs, const float* w,
                         uint32_t B, double* grad);
void   orc_adam(double* params, const double* grad, double* m, double* v, uint32_t n, uint32_t t,
                double lr, double b1, double b2, double eps);

#ifdef __cplusplus
}
#endif
#endif

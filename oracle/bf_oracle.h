/*
 * ORACLE — test infrastructure only. Nothing in the product links, includes or calls this.
 *
 * A plain-C CPU restatement of the reference's wavefront path-tracing hot path, used by tests/
 * as the parity checker, by __graft_entry__.smoke() as the smoke checker and by bench.py as the
 * CPU baseline (a std::thread-style worker pool over contiguous lane partitions, as the
 * reference's CpuWavefrontScheduler does). Every function cites the reference file:line it
 * restates; paths are relative to the reference repository root.
 *
 * Parity status per part is recorded in oracle/README.md.
 */
#ifndef BF_ORACLE_H
#define BF_ORACLE_H

#include "../include/blackframe_hip_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bfo_scene bfo_scene;

/* ---- sampling ----------------------------------------------------------------------------- */
uint64_t bfo_mix_bits(uint64_t value);
uint64_t bfo_indexed_bits(const bf_sample_stream_index* index, uint64_t dimension);
uint64_t bfo_sobol_bits(uint64_t sample_index, uint64_t sobol_dimension);
uint64_t bfo_owen_bits(const bf_sample_stream_index* index, uint64_t sobol_dimension);
float bfo_sample_1d(uint32_t sampler, const bf_sample_stream_index* index, uint64_t dimension);
uint64_t bfo_sobol_dimension_count(void);
/* Transport dimensions of bounce b (light selection, light u/v, bsdf component/u/v, medium x3,
 * roulette): returns 0, or 1 with zeros past the mapped range. */
int bfo_bounce_dimensions(uint64_t bounce_index, uint64_t out[10]);

/* ---- film sensor -------------------------------------------------------------------------- */
void bfo_cie_xyz(const float radiance[4], const float wavelengths[4], float pdf, float xyz[3]);
void bfo_xyz_to_rgb(const float xyz[3], float rgb[3]);
double bfo_cie_y_integral(void);
void bfo_cie_cmf64(float wavelength, double out[3]);
uint64_t bfo_cie_fingerprint(void);

/* ---- scene ---------------------------------------------------------------------------------- */
int bfo_scene_build(const bf_scene_desc* desc, bfo_scene** out);
void bfo_scene_free(bfo_scene* scene);
/* canonical BVH arrays: pass NULL to query counts */
int bfo_scene_bvh(const bfo_scene* scene, uint64_t counts[5], bf_bvh_blas* blases,
                  bf_bvh_node* blas_nodes, uint32_t* primitive_refs, bf_bvh_node* tlas_nodes,
                  bf_bvh_instance_ref* instance_refs);
uint32_t bfo_scene_tlas_root(const bfo_scene* scene);
/* mesh-area-light tables: per light (instance index, triangle count, inverse area) */
uint32_t bfo_scene_light_count(const bfo_scene* scene);
/* The scalar per-path integrator (Engine/Source/SceneMisPathLoop.cpp:632-1016): one path at a time,
 * brute-force scene queries (checked against the BVH; near-coplanar ties take the BVH's answer and
 * are counted per path in path_ties, optional), the light sample's shadow ray resolved at its
 * vertex. One thread. A query where brute force and BVH disagree otherwise fails the path (70). */
int bfo_render_scalar(const bfo_scene* scene, const bf_frame_desc* frame, const bf_transport_options* options,
                      uint64_t path_begin, uint64_t path_count, bf_path_result* paths, bf_transport_report* report,
                      uint32_t* path_ties);
/* Power-weighted light selection (Renderer/Source/LightSampler.cpp:116-249, 605-624): cdf receives
 * n + 1 boundaries from n four-lane spectral powers; 0 or a bf_status. */
int bfo_light_power_cdf(const float* powers, uint32_t n, float* cdf);
/* Binds the power CDF of [count][4] powers (registry order: punctual, then mesh lights) to the
 * scene's transport; powers == NULL returns it to uniform selection. */
int bfo_scene_set_light_powers(bfo_scene* scene, const float* powers, uint32_t count);
/* LightSampler::sample (:1023-1057) and ::probability (:1133-1151) over a CDF; 1 = ok. */
int bfo_light_cdf_sample(const float* cdf, uint32_t n, float u, uint32_t* index, float* probability);
int bfo_light_cdf_probability(const float* cdf, uint32_t n, uint32_t i, float* probability);
int bfo_scene_light(const bfo_scene* scene, uint32_t light, uint32_t* instance, uint32_t* triangle_count,
                    float* inverse_area, float* areas, float* probabilities, float* next_probabilities);
int bfo_triangle_area(const float vertices[9], float* area, float normal[3]);

/* ---- traversal -------------------------------------------------------------------------------
 * counters (optional, may be NULL): [0] canonical nodes visited, [1] triangle tests,
 * [2] instances entered, per ray. */
void bfo_closest_hit(const bfo_scene* scene, const bf_transport_ray* rays, uint32_t count,
                     bf_closest_hit_result* results, uint32_t* counters);
void bfo_occlusion(const bfo_scene* scene, const bf_transport_ray* rays, uint32_t count,
                   bf_occlusion_result* results, uint32_t* counters);
void bfo_closest_hit_brute_force(const bfo_scene* scene, const bf_transport_ray* rays,
                                 uint32_t count, bf_closest_hit_result* results);

/* ---- camera --------------------------------------------------------------------------------- */
int bfo_camera_ray(const bf_camera_desc* camera, uint32_t sampler, const bf_sample_stream_index* index,
                   bf_transport_ray* ray);
int bfo_camera_cone(const bf_camera_desc* camera, uint32_t sampler, const bf_sample_stream_index* index,
                    const bf_transport_ray* ray, bf_ray_cone* cone);

/* ---- transport ------------------------------------------------------------------------------
 * Renders paths [path_begin, path_begin + path_count) of the frame with the 7-queue CPU
 * wavefront. Outputs are optional (NULL). film_* are indexed [pixel*3 + channel] and accumulate
 * in path-slot order (sample-major), matching the reference's host film loop.
 * queue_log (optional, 2*64 entries): per bounce FNV-1a of the ray / shadow queue contents.
 * threads: worker count for each stage (contiguous partitions); 0 = 1. */
int bfo_render(const bfo_scene* scene, const bf_frame_desc* frame,
               const bf_transport_options* options, uint64_t path_begin, uint64_t path_count,
               uint32_t threads, bf_path_result* paths, float* film_float, double* film_compensated,
               uint32_t* film_counts, bf_transport_report* report, uint64_t* traversal_counters);

/* Forward render that also writes the adjoint tape of each path (frame path order): final
 * radiance [4 per path] and the bit mask of unoccluded light samples by path vertex. */
int bfo_render_taped(const bfo_scene* scene, const bf_frame_desc* frame, const bf_transport_options* options,
                     uint64_t path_begin, uint64_t path_count, uint32_t threads, float* film_float,
                     double* film_compensated, uint32_t* film_counts, float* tape_radiance, uint64_t* tape_visibility);
/* Forward render with the Russian-roulette survival probabilities of each path vertex recorded
 * into (BF_SURVIVAL_RECORD) or decided by (BF_SURVIVAL_REPLAY) `survival`, [global path * (diffuse
 * + glossy + specular limits + 1) + vertex]; tape buffers optional (both or neither). */
int bfo_render_survival(const bfo_scene* scene, const bf_frame_desc* frame, const bf_transport_options* options,
                        uint64_t path_begin, uint64_t path_count, uint32_t threads, uint32_t survival_mode, float* survival,
                        bf_path_result* paths, float* film_float, double* film_compensated, uint32_t* film_counts,
                        float* tape_radiance, uint64_t* tape_visibility);
/* The path-input transport (trace_cuda_wavefront_transport's contract): `count` paths start from
 * the given sample streams, rays, ray cones and path states (host arrays) instead of the camera and
 * end in the terminal path states, rays, cones and lane controls. */
int bfo_render_paths(const bfo_scene* scene, const bf_transport_options* options, uint64_t count, uint32_t threads,
                     const bf_sample_stream_index* samples, const bf_transport_ray* rays, const bf_ray_cone* cones,
                     const bf_path_state_lane* states, bf_path_state_lane* out_states, bf_transport_ray* out_rays,
                     bf_ray_cone* out_cones, bf_lane_control* out_controls, bf_transport_report* report);
/* Path-replay adjoint of paths [path_begin, path_begin + path_count): texel gradients of
 * dLoss/d(film RGB sums) = film_gradient, ACCUMULATED in binary64 into texture_gradients[k]
 * (layout of scene texture k; NULL entries skipped). */
int bfo_render_adjoint(const bfo_scene* scene, const bf_frame_desc* frame, const bf_transport_options* options,
                       uint64_t path_begin, uint64_t path_count, uint32_t threads, const float* tape_radiance,
                       const uint64_t* tape_visibility, const float* film_gradient, double* const* texture_gradients);

/* ---- closure library (radiance transport) ------------------------------------------------
 * allowed: direction mask (BF_LOBE_REFLECTION / TRANSMISSION bits) after depth filtering.
 * Return values are the lobe statuses: 0 ok, 1 invalid, 2 not representable, 3 outside support. */
int bfo_closure_eval(const bf_closure_record* closure, uint32_t allowed, const float wo[3], const float wi[3],
                     float value[4]);
int bfo_closure_pdf(const bf_closure_record* closure, uint32_t allowed, const float wo[3], const float wi[3],
                    float* pdf);
int bfo_closure_sample(const bf_closure_record* closure, uint32_t allowed, const float wo[3], float event, float u,
                       float v, float out[9], uint32_t* lobes);
int bfo_dielectric_fresnel(float cos_i, float eta_i, float eta_t, float* out);
int bfo_conductor_fresnel(float cos_i, float eta, float k, float* out);
int bfo_ggx_d(float ax, float ay, const float m[3], float* out);
int bfo_smith_g1(float ax, float ay, const float d[3], float* out);
int bfo_smith_g2(float ax, float ay, const float wo[3], const float wi[3], float* out);
int bfo_smith_lambda(float ax, float ay, const float d[3], float* out);
int bfo_vndf_pdf(float ax, float ay, const float wo[3], const float m[3], float* out);
int bfo_sample_vndf(float ax, float ay, const float wo[3], float u, float v, float m[3], float* pdf);
int bfo_ggx_widths_representable(float ax, float ay);
/* texture lookup (bilinear, repeat) and RGB-basis reflectance of blackframe_texture.h */
int bfo_texture_sample(uint32_t width, uint32_t height, uint32_t channels, const float* texels, float u, float v,
                       float out[3], uint32_t index[4], float weight[4]);
float bfo_texel_reflectance(const float* value, uint32_t channels, float wavelength);
/* adjoint derivatives: d ln f / d(alpha_x, alpha_y) of GGX closures, d f / d w per lane */
int bfo_closure_alpha_log_gradient(const bf_closure_record* closure, const float wo[3], const float wi[3], double out[2]);
int bfo_closure_weight_derivative(const bf_closure_record* closure, uint32_t allowed, const float wo[3], const float wi[3],
                                  double out[4]);
/* MIS weight of `sampled` against `competing` (same measure); heuristic 0 balance, 1 power. */
int bfo_mis_weight(uint32_t heuristic, float sampled, float competing, float* weight);
/* Russian roulette on a 4-lane throughput (updated in place when the path survives). */
int bfo_russian_roulette(float throughput[4], float eta_scale, float u, float minimum, float maximum, int* survived);
/* Four circularly stratified wavelengths in [360, 830) nm from one sample. */
void bfo_sample_wavelengths(float u, float wavelengths[4]);
float bfo_hypotf(float x, float y);
float bfo_atan2f(float y, float x);
void bfo_sincosf(float angle, float* sine, float* cosine);

/* Surface maps (CudaWavefrontStages.cu:2992-3673): the EWA filter of one channel of an image over
 * (du/dx, dv/dx, du/dy, dv/dy), the ray cone's (u, v) differentials, and a material's maps applied
 * to a shading normal (ns in and out). Statuses: 0 ok, 1 invalid, 2 numerical failure, 3 over the
 * texel budget. */
int bfo_map_filter(const bf_image_texture_desc* image, float u, float v, const float differentials[4], uint32_t channel,
                   uint32_t u_wrap, uint32_t v_wrap, uint32_t max_anisotropy, uint32_t max_visits, float* out);
int bfo_map_cone_differentials(const float cone[2], const float direction[3], const float ng[3], const float dpdu[3],
                               const float dpdv[3], float out[4]);
int bfo_map_apply(const bfo_scene* scene, uint32_t material, const float cone[2], const float direction[3],
                  const float uv[2], const float ng[3], const float dpdu[3], const float dpdv[3], float ns[3]);
float bfo_expf(float x);
float bfo_expm1f(float x);
float bfo_logf(float x);

#ifdef __cplusplus
}
#endif
#endif

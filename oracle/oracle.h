/*
 * oracle.h — CPU restatement (FP64, with an FP32 instantiation) of the reference
 * renderer's ray-tracing path: acceleration-structure traversal semantics, the
 * raygen / closest-hit / any-hit / miss shaders and their libraries
 * (crates/examples/gltf_viewer/shaders/RayTracing.*, shaders/lib/*.glsl).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the CPU reference that tests compare the HIP library in
 * rustracer_amd/csrc against; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline load it.  Nothing in the product path calls it.
 *
 * Pins (see tests/test_oracle_*.py): RNG against an independent numpy restatement
 * and committed known-answer vectors; BRDF against closed forms and numerical
 * integrals; traversal against brute force; the integrator against analytic scenes.
 */
#ifndef RT_ORACLE_H
#define RT_ORACLE_H

#include <stdint.h>
#include "../include/rt/rt.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One bottom-level geometry of the oracle scene (host-side rt_blas_input). */
typedef struct orc_blas {
    uint32_t first_vertex;   /* PrimInfo.v_offset */
    uint32_t index_offset;   /* PrimInfo.i_offset (elements) */
    uint32_t prim_count;
    uint32_t opaque;         /* geometry flag OPAQUE */
} orc_blas;

/* Host copies of the pipeline's buffers.  Texture data pointers are host
 * pointers; instances[i].acceleration_structure_reference is a BLAS index. */
typedef struct orc_scene_desc {
    const rt_vertex* vertices; uint64_t num_vertices;
    const uint32_t* indices; uint64_t num_indices;
    const rt_prim_info* prim_infos; uint32_t num_prim_infos;
    const rt_material_raw* materials; uint32_t num_materials;
    const rt_light* dlights; uint32_t num_dlights;
    const rt_light* plights; uint32_t num_plights;
    const rt_texture* textures; uint32_t num_textures;
    const orc_blas* blases; uint32_t num_blases;
    const rt_instance* instances; uint32_t num_instances;
    const rt_texture* skybox;  /* six square faces +x, -x, +y, -y, +z, -z, or NULL */
} orc_scene_desc;

typedef struct orc_stats {
    uint64_t primary_rays, extension_rays, shadow_rays;
    uint64_t node_visits, tri_tests, shade_events, anyhit_calls;
} orc_stats;

/* --- RNG (shaders/lib/Random.glsl) --- */
void orc_pcg4d(const uint32_t in[4], uint32_t out[4]);
/* n successive rand() draws from state (x,y,z,w); returns the final w. */
uint32_t orc_rand_sequence(const uint32_t state[4], uint32_t n, float* out);
uint32_t orc_tea16(uint32_t v0, uint32_t v1);
/* n successive RandomFloat() draws of the LCG; returns the final seed. */
uint32_t orc_lcg_sequence(uint32_t seed, uint32_t n, float* out);
/* Stream-B seed that replaces clockARB() (RayTracing.rgen:33,40). */
uint32_t orc_stream_b_seed(uint32_t random_seed, uint32_t frame_count);

/* --- BRDF (shaders/lib/PBR.glsl), all FP64 --- */
typedef struct orc_matbrdf {
    double base_color[3];
    double metallic, roughness, ior, transmission;
    double specular_factor;
    double specular_color_factor[3];
    uint32_t use_spec, front_face, volume, workflow_sg;
    double sg_specular[3];       /* used when workflow_sg */
    double attenuation_color[3];
    double attenuation_distance;
    /* outputs of matBuild (filled by orc_mat_build) */
    double F0[3], F90[3], c_diff[3];
} orc_matbrdf;

void orc_mat_build(orc_matbrdf* m);
void orc_eval_combined_brdf(const double N[3], const double L[3], const double V[3],
                            const orc_matbrdf* m, double out[3]);
void orc_brdf_probability(const orc_matbrdf* m, const double V[3], const double N[3],
                          double out[3]);
/* brdf_type: 1 diffuse, 2 specular, 3 transmission.  Returns needScatter. */
int orc_sample_indirect(const double u[2], const double shading_n[3], const double geo_n[3],
                        const double V[3], const orc_matbrdf* m, uint32_t brdf_type,
                        double dir[3], double weight[3]);
/* The same with Ray.volume_dis in and out (the transmission lobe's attenuation). */
int orc_sample_indirect_vd(const double u[2], const double shading_n[3], const double geo_n[3],
                           const double V[3], const orc_matbrdf* m, uint32_t brdf_type, double* volume_dis,
                           double dir[3], double weight[3]);
double orc_ggx_d(double alpha_squared, double n_dot_h);
double orc_smith_g2_height_correlated(double alpha, double n_dot_l, double n_dot_v);
double orc_smith_g1_ggx(double alpha, double n_dot_s, double alpha_sq, double n_dot_s_sq);
void orc_sample_ggx_vndf(const double Ve[3], double alpha, const double u[2], double out[3]);
void orc_sample_hemisphere(const double u[2], double out[3]);
void orc_rotation_to_z(const double v[3], double q[4]);
void orc_rotate_point(const double q[4], const double v[3], double out[3]);
double orc_convert_metallic(const double diffuse[3], const double specular[3], double max_specular);
double orc_frostbite_diffuse(double roughness, double n_dot_l, double n_dot_v, double l_dot_h);
double orc_light_intensity_luminance(const rt_light* light, double distance);
void orc_offset_ray(const float p[3], const float n[3], float out[3]);

/* --- texture sampling (LOD 0), host texture --- */
void orc_texture_sample(const rt_texture* tex, double u, double v, double out[4]);
/* cube map (six square faces +x, -x, +y, -y, +z, -z, one after another) at LOD 0 */
void orc_cube_sample(const rt_texture* cube, const double dir[3], double out[4]);

/* --- ray/triangle (Woop, Benthin & Wald 2013), FP64 --- */
int orc_intersect_triangle(const double o[3], const double d[3], const double v0[3],
                           const double v1[3], const double v2[3], double tmin, double tmax,
                           double out_tuv[3]);

/* --- scene --- */
void* orc_scene_create(const orc_scene_desc* desc);
void orc_scene_destroy(void* scene);
/* Detached evaluation for finite differences of the adjoint's derivative: BRDF
 * values (NEE, sampled weights) use `tex` in place of the scene's textures while
 * lobe choices, directions, densities and Russian roulette keep the scene's
 * textures.  n = 0 switches it off.  The arrays must outlive the renders. */
void orc_scene_set_eval_textures(void* scene, const rt_texture* tex, uint32_t n);
/* Detached evaluation of light intensities: the NEE term (rchit:347-348) uses
 * plight_intensity[i] for point light i, and the miss shader (rmiss:18-26)
 * dlight_intensity[i] for directional light i, while the RIS choice and weights keep
 * the scene's intensities.  NULL (or n = 0) switches either off.  Copied. */
void orc_scene_set_eval_lights(void* scene, const float* plight_intensity, uint32_t np,
                               const float* dlight_intensity, uint32_t nd);
/* rays: n x {ox,oy,oz,tmin,dx,dy,dz,tmax} (float).  hits: n x rt_hit (t,u,v in
 * float for layout compatibility) plus n x 3 doubles of (t,u,v) when hits_tuv is
 * non-NULL.  brute != 0 tests every triangle of every instance.  Without
 * RT_RAY_FLAG_OPAQUE, non-opaque geometry runs the any-hit alpha test; ray i draws
 * from the stream (i, 0, seed, 0). */
int orc_trace_rays(void* scene, const float* rays, uint32_t n, uint32_t ray_flags, uint32_t seed,
                   rt_hit* hits, double* hits_tuv, int brute, int threads);
/* Renders the n pixels xy[2i], xy[2i+1] of a width x height launch; out receives
 * n x 4 doubles: summed radiance and the payload's t after the pixel's last trace. */
int orc_render_pixels(void* scene, const rt_ubo* ubo, uint32_t width, uint32_t height,
                      const uint32_t* xy, uint32_t n, int precision, int threads, double* out,
                      orc_stats* stats);
/* Which pixels read which watched texels in the detached evaluation: renders the n
 * pixels xy (FP64) with the eval textures set (else -4) and writes (entry, i) pairs,
 * entry k being texel watch_texel[k] of texture watch_tex[k], for every pixel i whose
 * eval texture fetches read it (bilinear taps included).  Pixels that read none of
 * them do not depend on them.  Returns -3 when more than `capacity` pairs exist
 * (*n_out is then the number needed). */
int orc_texel_footprints(void* scene, const rt_ubo* ubo, uint32_t width, uint32_t height, const uint32_t* xy,
                         uint32_t n, const uint32_t* watch_tex, const uint32_t* watch_texel, uint32_t n_watch,
                         uint32_t* out_pairs, uint64_t capacity, uint64_t* n_out, int threads);
/* Primary hits of sample 0 (camera ray and jitter as RayTracing.rgen computes them). */
int orc_trace_primary(void* scene, const rt_ubo* ubo, uint32_t width, uint32_t height,
                      rt_hit* hits, double* hits_tuv, int threads);
/* Renders pixels [x0,x1) x [y0,y1) of a width x height launch into accum
 * (width*height*4 doubles, updated in place per the accumulation rule).
 * precision: 64 or 32. */
int orc_render(void* scene, const rt_ubo* ubo, uint32_t width, uint32_t height,
               uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1, int precision,
               int threads, double* accum, orc_stats* stats);

/* --- skinning (AnimationCompute.comp:14-38), fp64; returns the count of vertices
 * left unskinned because their skin index or a joint is out of range --- */
uint32_t orc_skin_vertices(const rt_vertex* in, rt_vertex* out, uint32_t n, const float* skins, uint32_t num_skins);

/* --- display image (RayTracing.rgen:132-166), fp32 --- */
float orc_tone_map(float c, uint32_t mode);
void orc_heatmap(float t, float out[3]);
/* n pixels of an rgba32f accumulation image -> packed 8-bit pixels (R,G,B,A bytes,
 * or B,G,R,A when bgra).  last_t: payload t per pixel (DISTANCE); ticks: shader
 * clock ticks per pixel (HEAT); either may be NULL (read as 0). */
void orc_output_image(const float* accum, uint32_t n, const float* last_t, const uint64_t* ticks, const rt_ubo* ubo,
                      int bgra, uint32_t* out);

#ifdef __cplusplus
}
#endif

#endif

/*
 * cg_oracle.h — ORACLE: test infrastructure only.
 *
 * A scalar CPU restatement of the reference's spectral path-tracing core:
 * the WGSL compute pass (src/wgpu/shader/compute.wesl and the compute/ modules) and the
 * host scene preparation (src/pathtracing/ headers).  It exists to check the
 * product (the HIP library under crystalgraphics_amd/) and is only ever
 * loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 * Nothing in the product links, loads or calls it.
 *
 * Float semantics are declared, not inherited from Tint/Dawn:
 *   - every + - * / sqrt is one IEEE-754 binary32 operation, evaluated left to
 *     right as written in the WGSL, with no FMA contraction;
 *   - min/max/clamp ignore a NaN operand (IEEE minNum/maxNum, fminf/fmaxf);
 *   - f32 -> u32 saturates and maps NaN / negatives to 0;
 *   - runtime array indices are clamped to the array (WebGPU robustness);
 *   - pow(x, 2) = x*x and pow(x, 4) = (x*x)*(x*x);
 *   - tan(acos(c))^2 in the GGX masking terms = max(0, 1 - c^2) / c^2;
 *   - sin/cos of the sampling angles: Cody-Waite reduction and Taylor
 *     polynomials (orc_sincos; about 1 ulp, inside WGSL's 2^-11 bound);
 *   - exp/acos/atan2 are the C library's; the sRGB power pow(x, f32(1/2.4))
 *     is the double-precision power rounded once to binary32 (so the device,
 *     which computes it the same way, matches it bit for bit);
 *   - material texture lookups: repeat-wrap in normalised coordinates, texel
 *     centres at (i + 1/2) / size, texels k / 255, and the blend
 *     lerp(lerp(T00, T10, a), lerp(T01, T11, a), b) with
 *     lerp(p, q, w) = p + w * (q - p) (orc_texture_sample).
 */
#ifndef CG_ORACLE_H_
#define CG_ORACLE_H_

#include <stdint.h>

#include "../include/cg_scene.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_scene orc_scene;

/* Scene preparation (src/pathtracing/scene_data.h:38-54): BLAS per primitive
 * object, TLAS, instances, emissive buffer (with its emissive-texture
 * lookups), alias table, material texture array, fallback sky. */
orc_scene* orc_scene_build(const cg_vertex* vertices, uint32_t vertex_count,
                           const uint32_t* indices, uint32_t index_count,
                           const cg_material* materials, uint32_t material_count,
                           const cg_primitive* prims, uint32_t prim_count,
                           const cg_texture* textures, uint32_t texture_count,
                           const float* env_rgba32f, uint32_t env_w, uint32_t env_h);
/* Take a SceneData built elsewhere (copied). */
orc_scene* orc_scene_wrap(const cg_scene_data* sd);
void orc_scene_free(orc_scene* s);
/* Pointers into the oracle scene's storage. */
void orc_scene_get(const orc_scene* s, cg_scene_data* out);

/* Closest hit per ray (bvh.wesl:52-186).  second_t (nullable) receives the
 * second-smallest hit distance among distinct triangles (1e6 if none), so a
 * caller can exclude ties.  counts (nullable) receives per ray
 * {TLAS nodes, BLAS nodes, triangles, instances} visited.  cull != 0 skips
 * boxes entered beyond the current closest hit (same result, faster). */
void orc_trace_closest(const orc_scene* s, const cg_ray* rays, uint32_t n,
                       cg_hit* hits, float* uv, float* second_t,
                       uint32_t* counts, int cull, int threads);

/* The compute pass over pixel rows [y0, y1) (compute.wesl:32-62).  history is
 * the full W*H buffer, read when iter_count > 0 and written for the rows.
 * rgba8 (nullable, W*H*4), path_hash (nullable, W*H) receives a hash of the
 * pixel's sequence of hit ids, shadow outcomes (each with the emitter its
 * light sample picked) and sky texels, ray_count
 * (nullable, W*H)
 * its closest-hit queries.  Returns 0, or -1 on an unsupported scene. */
int orc_render(const orc_scene* s, const cg_conf* conf, const cg_camera* cam,
               uint32_t iter_count, cg_history_pixel* history, uint8_t* rgba8,
               uint32_t* path_hash, uint32_t* ray_count,
               uint32_t y0, uint32_t y1, int threads, int cull);

/* orc_render with a choice of RNG mode: 0 the reference's one stream per
 * pixel (orc_render), 1 one stream per sample -- sample i seeded
 * orc_sample_seed(b, i) from the pixel's base seed b, XYZ summed in sample
 * order, carried seed orc_sample_carry(b), path hash the fold of the
 * samples' hashes in sample order. */
int orc_render_mode(const orc_scene* s, const cg_conf* conf, const cg_camera* cam,
                    uint32_t iter_count, cg_history_pixel* history, uint8_t* rgba8,
                    uint32_t* path_hash, uint32_t* ray_count,
                    uint32_t y0, uint32_t y1, int threads, int cull, int rng_mode);
uint32_t orc_sample_seed(uint32_t base, uint32_t i);
uint32_t orc_sample_carry(uint32_t base);

/* Unit-level pieces, for known-answer and closed-form checks. */
/* Material texture lookup (bilinear, repeat, level 0) in an array of `layers`
 * w x h RGBA8 images; out receives rgba. */
void orc_texture_sample(const uint8_t* rgba8, uint32_t w, uint32_t h, uint32_t layers,
                        int32_t layer, float u, float v, float* out);
/* The emissive buffer's nearest lookup at (u, v) under t's wrap modes. */
void orc_emissive_texel(const cg_texture* t, float u, float v, float* rgb);
void orc_sincos(float x, float* s, float* c);
uint32_t orc_rng_step(uint32_t state);
float orc_rand_float(uint32_t* state);
float orc_erf(float x);
void orc_xyz_match(float lambda, float* xyz);
float orc_xyz_match_sum_integral(float lo, float hi);
void orc_sample_wavelength(float u, float* lambda, float* prob);
void orc_illuminate(float lambda, float radiance, float* xyz);
float orc_smits(int curve, float lambda); /* 0 white 1 red 2 green 3 blue 4 magenta 5 cyan 6 yellow */
float orc_rgb_to_reflectance(const float* rgb, float lambda);
float orc_rgb_to_radiance(const float* rgb, float lambda);
float orc_dispersed_ior(float nd, float dispersion, float lambda);
float orc_fresnel(const float* dir, const float* normal, float ior_re, float ior_im);
void orc_alias_build(const double* weights, uint32_t n, cg_alias_entry* out);
uint32_t orc_alias_pick(const cg_alias_entry* table, uint32_t n, float u1, float u2);
void orc_srgb(const float* xyz, float* rgb);
void orc_camera_ray(const cg_camera* cam, uint32_t w, uint32_t h,
                    uint32_t px, uint32_t py, cg_ray* out);
int orc_intersect_triangle(const cg_ray* ray, const float* v0, const float* v1,
                           const float* v2, float* tuv);
int orc_intersect_aabb(const cg_ray* ray, const float* lb, const float* ub);
/* Sampling of one lobe at a surface for unit tests: model 1 diffuse,
 * 3 rough reflect, 5 rough transmit.  Returns the pdf, writes the direction. */
float orc_sample_lobe(int model, const float* curr_dir, const float* shading_norm,
                      float alpha, float ior_re, float ior_im, float transmission,
                      uint32_t* rng_state, float* next_dir);

/* A decoded material at one wavelength (material.wesl:60-67). */
typedef struct orc_mat {
  float diffuse, alpha, ior_re, ior_im, transmission;
  uint32_t volume;
} orc_mat;
/* One lobe at a surface with shading normal sn (= geometric normal here):
 * model 1 diffuse, 2 specular reflection, 3 rough reflection, 4 specular
 * transmission, 5 rough transmission.  orc_lobe_sample draws a direction as
 * sampleModels does and returns its pdf; orc_lobe_eval returns the lobe's
 * pdf at next_dir and writes its dist (f times the projected cosine). */
float orc_lobe_sample(int model, const float* curr_dir, const float* shading_norm, const orc_mat* m,
                      uint32_t* rng_state, float* next_dir);
float orc_lobe_eval(int model, const float* curr_dir, const float* shading_norm, const orc_mat* m,
                    const float* next_dir, float* dist);
/* GGX pieces with the reference's alpha (rough_reflection.wesl:120-139). */
float orc_ggx_d(const float* micro, const float* macro, float alpha);
float orc_ggx_g1(float c, float alpha);
float orc_ggx_g2(float c1, float c2, float alpha);
/* The visible-normal sampler (rough_reflection.wesl:52-103). */
void orc_sample_micro_normal(const float* curr_dir, const float* shading_norm, float alpha,
                             uint32_t* rng_state, float* micro);
void orc_tangent_space(const float* n, float* t, float* b);  /* geometric.wesl:28-33 */
void orc_reflect(const float* d, const float* n, float* out);  /* WGSL reflect */
void orc_refract(const float* d, const float* n, float eta, float* out);  /* WGSL refract */
/* Lobe counts (sample.wesl:110-246) at lod depth `depth` for conf's
 * trace depth and caps: out = {diffuse, specular reflect, rough reflect,
 * specular transmit, rough transmit, light}. */
void orc_sample_counts(const orc_mat* m, int can_sample, int depth, const cg_conf* conf, uint32_t* out);
/* balancePdfDenominator (sample.wesl:452-501) at a surface point of scene s;
 * model / model_pdf name the sampled lobe (0: none), sampled_light_pdf >= 0
 * replaces the light pdf. */
float orc_balance_denominator(const orc_scene* s, const float* curr_dir, const float* pos,
                              const float* shading_norm, const orc_mat* m, const uint32_t* counts,
                              int model, float model_pdf, float sampled_light_pdf,
                              const float* next_dir);
/* lightPdf (pdf/emission.wesl:120-149) from pos along next_dir. */
float orc_light_pdf(const orc_scene* s, const float* pos, const float* next_dir);
/* emissionSample (pdf/emission.wesl:28-96) from pos: returns the solid-angle
 * pdf; writes the direction, the distance to the light point, the radiance
 * at lambda and (nullable) the picked emitter. */
float orc_emission_sample(const orc_scene* s, float lambda, const float* pos, uint32_t* rng_state,
                          float* next_dir, float* expected_dist, float* radiance, uint32_t* prim);
/* spawnSurfaceRay's origin (spawn_surface_ray.wesl:6-15). */
void orc_spawn(const float* pos, const float* geom_norm, const float* next_dir, float* out_pos);

#ifdef __cplusplus
}
#endif

#endif /* CG_ORACLE_H_ */

/*
 * Oracle of the MI355X path tracer: a CPU restatement of the reference
 * renderer's per-pixel integrator and the stages it calls.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load liboracle.so, and only as the checker
 * (or the timed CPU baseline); the product never calls it.
 *
 * Each function follows a reference file (paths relative to Part_2/src of
 * Neroued/PathTracingTutorial); see oracle.cpp for file:line citations.
 * Render functions are instantiated in float (the reference's precision) and
 * in double (the converged FP64 oracle image).
 */
#ifndef PT_ORACLE_H
#define PT_ORACLE_H

#include <stdint.h>

#include "pt_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_scene oracle_scene;

/* Builds the reference BVH (bins/leaf_threshold; 24/4 are the reference's
 * constants) and the light tables over the given host arrays. */
oracle_scene* oracle_scene_create(const pt_vertex* v, uint32_t nv, const pt_face* f, uint32_t nf,
                                  const pt_material* m, uint32_t nm, const float* env_rgba, int env_w, int env_h,
                                  const pt_punctual_light* lights, uint32_t nl, int bins, int leaf_threshold);
/* Adds a texture (RGBA float, GL wrap enums, linear=1 skips the pow 2.2
 * colour decode); returns its index. */
int oracle_scene_add_texture(oracle_scene* s, const float* rgba, int w, int h, int wrap_s, int wrap_t, int linear);
void oracle_scene_destroy(oracle_scene* s);

/* BVH in the reference's depth-first layout: 8 floats per node (data words
 * as uint32 bits); face_order[i] = original face of BVH slot i. */
uint32_t oracle_bvh_num_nodes(const oracle_scene* s);
void oracle_bvh_get(const oracle_scene* s, float* nodes8, uint32_t* face_order);

/* Closest hit (face = original face id or -1, t = FLT_MAX on miss) or any hit. */
void oracle_trace(const oracle_scene* s, const float* rays8, int n, int any_hit, float* t, int32_t* face,
                  float* bary);

/* Light tables (light_sampler.cpp) of the scene: counts, then copies. */
void oracle_light_tables(const oracle_scene* s, uint32_t* n_emissive, uint32_t* n_env, float* tri_pmf,
                         float* tri_alias, float* env_pmf, float* env_alias);

/* Alias table of arbitrary weights (light_sampler.cpp:7-55); returns n or 0. */
uint32_t oracle_alias_table(const float* w, uint32_t n, float* prob, uint32_t* alias);

/* Samplers: raw 32-bit value and float of dimension dim. */
uint32_t oracle_sobol_u32(int px, int py, uint32_t index, uint32_t dim, uint32_t seed);
float oracle_sobol_float(int px, int py, uint32_t index, uint32_t dim, uint32_t seed);
/* n draws of the PCG sampler of pixel (x, y), frame, sample. */
void oracle_pcg(int x, int y, uint32_t frame, uint32_t sample, uint32_t seed, int n, uint32_t* u32, float* f);

/* Camera frame (camera/camera.h:36-51). */
void oracle_camera_frame(const pt_camera* cam, pt_camera_frame* out);

/* Disney BSDF of a material (fetch + clamps, no textures) at directions in
 * world space with shading normal n: f*NoL (or f for transmission) and pdf.
 * fp64 selects double arithmetic. */
void oracle_bsdf_eval(const pt_material* m, const double wo[3], const double wi[3], const double n[3], int entering,
                      int fp64, double f[3], double* pdf);
/* sampleDisneyBsdf with the three uniform numbers u[0..2] (choice, then the
 * two warp numbers in argument order).  Returns valid; wi, weight, pdf, delta. */
int oracle_bsdf_sample(const pt_material* m, const double wo[3], const double n[3], int entering, const float u[3],
                       double wi[3], double weight[3], double* pdf, int* delta);

/* Microfacet and sampling pieces (materials/microfacet.cuh, math/sampling.h,
 * math/frame.h) in float for the inputs in[12] = {noh, hx, hy, alpha, ax, ay,
 * nov, nol, cos, eta, u1, u2}: out[27] = D, D_aniso, G1, G2, Schlick(0.04),
 * Schlick(f0) x3, F_dielectric, VNDF half vector x3, its pdf, refract ok +
 * direction x3 (normal +z), cosine sample x3, power heuristic(nov, nol),
 * Frame(n) toLocal(wo) x3 and toWorld(wo) x3. */
void oracle_microfacet(const float* in, const float* f0, const float* wo, const float* n, float* out);
/* Equirect mapping and env pdf (envmap.h:15-32, light_pdf.cuh:12-27). */
void oracle_dir_to_uv(const float d[3], float uv[2]);
void oracle_uv_to_dir(const float uv[2], float d[3]);
float oracle_env_pdf(const oracle_scene* s, const float d[3]);
/* emissiveTrianglePdf (light_pdf.cuh:29-54), FP64: solid-angle pdf of hitting
 * primitive prim at distance dist along direction d. */
double oracle_emissive_pdf(const oracle_scene* s, int prim, double dist, const double d[3]);
/* sampleAreaLight (env = 0, direct_lighting.cuh:75-121) or sampleEnvLight
 * (env = 1, :123-154), FP64, at point p with normal n, drawing from the PCG
 * sequence of (x, y, frame, sample); returns 1 and the direction, radiance
 * and solid-angle pdf of a valid sample, else 0. */
int oracle_sample_light(const oracle_scene* s, int env, const double p[3], const double n[3], int x, int y,
                        uint32_t frame, uint32_t sample, double wi[3], double L[3], double* pdf);
/* Bilinear texture fetch with CUDA normalized-coordinate semantics. */
void oracle_tex2d(const float* rgba, int w, int h, int wrap_s, int wrap_t, float u, float v, float out[4]);

/* fetchMaterial (materials/disney_bsdf.cuh:53-122) of material m with the
 * scene's textures at texture coordinates (u, v), in float: out[20] = base
 * rgb, metallic, roughness, specTrans, ior, clearcoat, clearcoatRoughness,
 * sheen, sheenTint, subsurface, anisotropic, sheenColor rgb, emissive rgb,
 * specularFactor. */
void oracle_fetch_material(const oracle_scene* s, const pt_material* m, float u, float v, float out[20]);
/* The shading point the integrator builds at the closest hit of ray (o, d)
 * (path.cu:86-103), in float: out[16] = 1, t, P xyz, interpolated normal xyz,
 * the normal on the ray's side (double-sided flip) xyz, the shading normal
 * after the normal map xyz, entering, original face id.  Returns 0 (out all
 * zero) on a miss. */
int oracle_surface(const oracle_scene* s, const float o[3], const float d[3], float out[16]);

/* Renders pixels [x0,x1) x [y0,y1) of one batch (frame) of spf samples per
 * pixel, as kernelPathTrace does (integrators/path.cu:209-242): out gets the
 * batch mean RGB per pixel (row-major over the rectangle, 3 floats).
 * counters (may be null) gets {closest rays, shadow rays, paths}.
 * threads <= 0 uses all hardware threads.  fp64 = 1 renders in double. */
void oracle_render(const oracle_scene* s, const pt_camera* cam, int spf, int max_depth, int sampler, uint32_t seed,
                   uint32_t frame, int x0, int x1, int y0, int y1, int threads, int fp64, float* out,
                   uint64_t* counters);

/* The compiler command line this library was built with. */
const char* oracle_build_flags(void);

/* Gradient of  sum_pixels weights . I_batch  (I_batch = the batch mean image
 * of `frame`, weights W*H*3) with respect to nparams scalars, by central
 * differences with step h of the detached estimator in double precision:
 * every path is traced once at the current parameters (float, the
 * integrator's draws and decisions) and recorded; the recorded paths are then
 * evaluated with each parameter at +h and -h (directions, pdfs, MIS weights
 * and Russian-roulette factors kept).  kind: 0 material a baseColor channel c,
 * 1 material a roughness, 2 material a metallic, 3 texture a texel b channel
 * c, 4 environment texel b channel c.  value (may be null) gets the weighted
 * sum at the current parameters.  Parameters are floats: the difference is
 * divided by the step between the two rounded values actually used. */
void oracle_adjoint_fd(const oracle_scene* s, const pt_camera* cam, int spf, int max_depth, int sampler, uint32_t seed,
                       uint32_t frame, const float* weights, int nparams, const int* kind, const int* a, const int* b,
                       const int* c, double h, int threads, double* grad, double* value);

/* The host reference adjoint: the gradient of  sum_pixels weights . I_batch
 * (the same batch and paths as oracle_adjoint_fd) with respect to every
 * parameter at once, by reverse-mode differentiation in double of the
 * detached estimator of each recorded path.  g_mat (nm x 5: base colour rgb,
 * roughness, metallic), g_tex (the texel channels of every texture, in
 * texture order, 4 per texel), g_env (4 per environment texel); any may be
 * null.  Pixels whose weights are all zero are not traced. */
void oracle_adjoint(const oracle_scene* s, const pt_camera* cam, int spf, int max_depth, int sampler, uint32_t seed,
                    uint32_t frame, const float* weights, int threads, double* g_mat, double* g_tex, double* g_env,
                    double* value);

#ifdef __cplusplus
}
#endif
#endif

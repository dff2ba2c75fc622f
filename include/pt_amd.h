/*
 * pt_amd.h — C ABI of the MI355X path tracer (libptamd.so).
 *
 * Plain C: no HIP or torch types, only pointers, sizes and POD structs.
 * Every entry point returns a pt_status; pt_last_error() gives the message.
 * Handles (pt_scene, pt_film) own their device memory.  The library is
 * thread-compatible (one thread per handle at a time), not thread-safe.
 *
 * Each declaration cites the interface of the reference renderer
 * (Neroued/PathTracingTutorial, paths relative to Part_2/src) that it replaces.
 */
#ifndef PT_AMD_H
#define PT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PT_ABI_VERSION 1

typedef int pt_status;
enum {
    PT_OK = 0,
    PT_ERR_INVALID_ARG = 1,
    PT_ERR_IO = 2,
    PT_ERR_PARSE = 3,
    PT_ERR_DEVICE = 4,
    PT_ERR_STATE = 5,
    PT_ERR_UNSUPPORTED = 6,
    PT_ERR_OUT_OF_MEMORY = 7,
    PT_ERR_LIMIT = 8
};

/* ------------------------------------------------------------------------ */
/* Data contract: byte layouts identical to the reference's host structs so  */
/* that HostScene arrays convert by memcpy.                                   */
/* ------------------------------------------------------------------------ */

/* vec3 with the reference's alignas(16) padding (math/vecmath.h:63). */
typedef struct pt_vec3p { float x, y, z, _pad; } pt_vec3p;

/* Vertex, 64 B (geometry/vertex.h:10-15). */
typedef struct pt_vertex {
    pt_vec3p position;
    pt_vec3p normal;
    float u, v;
    float _pad[2];
    float tangent[4]; /* xyz direction, w handedness sign */
} pt_vertex;

/* TriangleFace, 16 B (geometry/vertex.h:17-20). */
typedef struct pt_face { uint32_t v0, v1, v2; int32_t material_id; } pt_face;

/* Material, 256 B (materials/material.h:8-81), field for field. */
typedef struct pt_material {
    pt_vec3p base_color;
    float metallic;
    float roughness, specular_tint, anisotropic, sheen;
    float sheen_tint, clearcoat, clearcoat_roughness, spec_trans;
    float ior, subsurface, specular_factor, specular_color_r;
    float specular_color_g, specular_color_b, anisotropy_rotation, normal_scale;
    float _pad_e[3];
    pt_vec3p emissive;
    float emissive_strength;
    float alpha_cutoff;
    int32_t alpha_mode; /* 0 opaque, 1 mask, 2 blend */
    float base_alpha;
    int32_t double_sided;
    float sheen_color_r, sheen_color_g, sheen_color_b, occlusion_strength;
    float thickness_factor, attenuation_distance, att_color_r, att_color_g;
    float att_color_b;
    float uv_offset_x, uv_offset_y, uv_scale_x;
    float uv_scale_y, uv_rotation, _pad1, _pad2;
    int32_t base_color_tex_id, metallic_roughness_tex_id, normal_tex_id, emissive_tex_id;
    int32_t occlusion_tex_id, clearcoat_tex_id, clearcoat_roughness_tex_id, clearcoat_normal_tex_id;
    int32_t transmission_tex_id, sheen_color_tex_id, sheen_roughness_tex_id, _pad3;
    int32_t _pad_tail[3];
} pt_material;

/* PunctualLight, 112 B (lights/punctual_light.h:14-25). type: 0 point, 1 directional, 2 spot. */
typedef struct pt_punctual_light {
    pt_vec3p position;
    float range;
    float _pad0[3];
    pt_vec3p color;
    float intensity;
    float _pad1[3];
    pt_vec3p direction;
    float inner_cone, outer_cone; /* cosines */
    int32_t type;
    float _pad2[2];
    float _tail[3];
} pt_punctual_light;

/* Camera, 80 B (camera/camera.h:12-21). fov is vertical, in degrees. */
typedef struct pt_camera {
    pt_vec3p position;
    pt_vec3p look_at;
    pt_vec3p up;
    float fov, aperture, focal_distance;
    int32_t width, height;
    float _pad[3];
} pt_camera;

/* CameraFrame as built by buildCameraFrame (camera/camera.h:25-51). */
typedef struct pt_camera_frame {
    float position[3], forward[3], right[3], up[3];
    float tan_half_fov, aspect, aperture, focal_distance;
} pt_camera_frame;

/* A texture: RGBA float texels, row 0 first; wrap_s/t are GL enums
 * (10497 repeat, 33071 clamp-to-edge, 33648 mirrored repeat) as in
 * io/image_io.h:8-17.  linear=0 applies the reference's pow(x, 2.2) decode to
 * colour fetches (materials/disney_bsdf.cuh:81-83); linear=1 skips it. */
typedef struct pt_texture_desc {
    const float* rgba;
    int32_t width, height;
    int32_t wrap_s, wrap_t;
    int32_t linear;
    int32_t _pad;
} pt_texture_desc;

/* Render settings (scene/scene.h:16-22 RenderSettings) plus a global seed.
 * sampler: 0 PCG, 1 Sobol.  seed 0 reproduces the reference's sequences. */
typedef struct pt_render_settings {
    int32_t sample_per_frame;
    int32_t target_samples;
    int32_t max_depth;
    int32_t sampler;
    uint32_t seed;
    int32_t _pad[3];
} pt_render_settings;

/* Host scene descriptor: the fields of HostScene (scene/scene.h:24-35). */
typedef struct pt_host_scene_desc {
    const pt_vertex* vertices; uint32_t num_vertices;
    const pt_face* faces; uint32_t num_faces;
    const pt_material* materials; uint32_t num_materials;
    const pt_texture_desc* textures; uint32_t num_textures;
    const pt_punctual_light* lights; uint32_t num_lights;
    const pt_camera* camera;
    /* environment map, RGBA float (channel 4 as device_scene.cu:19 requires); NULL for none */
    const float* env_rgba; int32_t env_width, env_height;
    const pt_render_settings* settings;
} pt_host_scene_desc;

/* BVH build options.  Reference settings (accel/bvh_builder.cpp:10-13):
 * bins 24, leaf_threshold 4 (leaf when count < 4), max_leaf 0 (no cap on the
 * SAH early-out leaf).  max_leaf > 0 forbids SAH leaves above that size. */
typedef struct pt_bvh_options {
    int32_t bins;
    int32_t leaf_threshold;
    int32_t max_leaf;
    int32_t builder; /* 0 = GPU binned SAH (default) */
} pt_bvh_options;

typedef struct pt_scene_info {
    uint32_t num_vertices, num_faces, num_materials, num_textures;
    uint32_t num_lights, num_emissive_triangles, num_bvh_nodes, bvh_depth;
    int32_t env_width, env_height;
    int32_t width, height;
    int32_t sample_per_frame, target_samples, max_depth, sampler;
    uint32_t seed, frame_count;
    float bvh_build_ms;
    uint32_t max_leaf_size;
} pt_scene_info;

/* Ray and hit counters of the last counted render (rays split as in-kernel). */
typedef struct pt_counters {
    uint64_t closest_rays;   /* camera + extension rays */
    uint64_t shadow_rays;    /* any-hit rays */
    uint64_t node_visits;    /* 32-B units of node data fetched (a binary node record; a 4-wide node is 2) */
    uint64_t tri_tests;      /* triangle intersection tests */
    uint64_t paths;
    uint64_t stack_overflows;
    uint64_t shadow_node_visits; /* the any-hit (shadow ray) share of node_visits */
    uint64_t shadow_tri_tests;   /* the any-hit share of tri_tests */
    /* events of the wavefront integrator's shading kernel (zero for the
     * megakernel), from which its algorithmic bytes follow: */
    uint64_t shade_items;        /* closest-hit results shaded */
    uint64_t shade_env_lookups;  /* escaped rays that read the environment */
    uint64_t shade_hits;         /* surface hits (interpolated surface + material read) */
    uint64_t shade_light_samples;  /* shading points that drew the light samples (NEE) */
    uint64_t shade_shadow_entries; /* shadow-queue entries written */
    uint64_t shade_continuations;  /* rays queued for the next bounce */
    uint64_t shade_radiance_adds;  /* read-modify-writes of a path's radiance sum */
} pt_counters;

typedef struct pt_scene pt_scene;
typedef struct pt_film pt_film;

enum { PT_INTEGRATOR_PATH = 0, PT_INTEGRATOR_WAVEFRONT = 1 };

/* Launch configuration of the wavefront integrator (no reference
 * counterpart: the reference has only the per-pixel megakernel,
 * integrators/path.cu:209-255).  Every field changes speed only, never the
 * estimate: the estimator and its random numbers are the megakernel's (with
 * bvh_width 4 a ray whose closest hit is an exact-distance tie between
 * triangles may take the other one). */
typedef struct pt_wavefront_config {
    int32_t chunk_spp;        /* paths in flight = chunk_spp x all jobs' pixels, at most 2^28 (default 64) */
    int32_t extend_block;     /* threads per block of the traversal kernels: 64, 128 or 256 */
    int32_t shade_block;      /* threads per block of the shading kernel: 64, 128 or 256 */
    int32_t waves_per_simd;   /* persistent grid: waves per SIMD, 0 = occupancy limit */
    int32_t sort_by_material; /* counting-sort each bounce's shading queue by material */
    int32_t use_graph;        /* replay one captured hipGraph per chunk */
    int32_t time_stages;      /* bracket every launch with HIP events (no graph): pt_scene_wavefront_stats */
    int32_t shade_waves;      /* shading kernel variant built for >= this many waves per SIMD: 0, 3 or 4 */
    int32_t stack_lds;        /* traversal stack entries held in LDS, 16 or 32 (the rest in device memory) */
    int32_t grab_extend;      /* queue items a wave takes per atomic in the closest-hit kernel: 64..1024, power of 2 */
    int32_t grab_shade;       /* the same for the shading kernel */
    int32_t grab_shadow;      /* the same for the shadow-ray kernels */
    int32_t bvh_width;        /* 4: traverse the 4-wide BVH with quantized child boxes (built with the
                                 binary one; the same closest hits up to exact-distance ties), 2: the
                                 binary BVH, children entered by slab entry distance (default 4) */
} pt_wavefront_config;

/* GPU time (ms) and launch count per stage of the last wavefront batch
 * rendered with time_stages set: 0 camera rays, 1 closest-hit traversal,
 * 2 material sort, 3 shading, 4 shadow-ray traversal, 5 control (queue
 * resets, per-chunk finalize, film). */
typedef struct pt_wavefront_stats {
    float stage_ms[6];
    uint32_t launches[6];
} pt_wavefront_stats;

/* ------------------------------------------------------------------------ */
/* Library                                                                    */
/* ------------------------------------------------------------------------ */
int pt_abi_version(void);
/* Message of the last failed call on this thread.  Replaces the reference's
 * print-and-exit CUDA_CHECK (cuda/check.h:6-24). */
const char* pt_last_error(void);
/* Selects and initialises the HIP device. */
pt_status pt_init(int device);
/* Struct sizes of the data contract, for binding checks: fills out[0..7] with
 * sizeof pt_vertex, pt_face, pt_material, pt_punctual_light, pt_camera,
 * pt_camera_frame, pt_host_scene_desc, pt_counters. */
pt_status pt_struct_sizes(uint32_t* out, int n);

/* ------------------------------------------------------------------------ */
/* Scene                                                                      */
/* ------------------------------------------------------------------------ */
/* Loads a scene by extension (scene/scene_loader.cpp:278-292 SceneLoader::load;
 * scene/scene_manager.cpp:11 SceneManager::load): .json (the same schema and
 * defaults, including "gltf" objects with transforms and materialOverrides,
 * scene_loader.cpp:162-194) or a glTF 2.0 file, .gltf or .glb
 * (io/gltf_loader.cpp:583-732: its camera, or one fitted to the bounds; its
 * KHR_lights_punctual lights, or two default lights when nothing emits). */
pt_status pt_scene_load(const char* path, pt_scene** out);
/* Builds a scene from host arrays (HostScene, scene/scene.h:24-35). */
pt_status pt_scene_create(const pt_host_scene_desc* desc, pt_scene** out);
/* Procedural workloads: 0 = Cornell box (36 tris, 512x512), 1 = 64 displaced
 * icospheres (1,310,720 tris + ground, 2048x1024 sky+sun env, 1920x1080),
 * 2 = 8 small icospheres (tests), 3 = differentiable Cornell box (8
 * materials, 256x256 albedo texture on the back wall, 512x256 environment,
 * 256x256, 64 spp, depth 6; the seeded values are a ground truth), 4 = one
 * job of the batched workload (icospheres of 1e3..5e5 triangles, 128^2..1024^2
 * pixels, 4..64 spp, depth 12, 512x256 sky; all drawn from the seed), 5 =
 * the multi-view differentiable scene (2,048,000 icosphere triangles, 16
 * materials, 1024x1024 albedo texture on the ground, 1024x512 sky; 512x512,
 * 32 spp, depth 10).
 * No reference counterpart; seeds fix every value. */
pt_status pt_scene_generate(int which, uint64_t seed, pt_scene** out);
void pt_scene_destroy(pt_scene* scene);
pt_status pt_scene_get_info(const pt_scene* scene, pt_scene_info* info);

/* SceneManager::buildAccel (scene/scene_manager.cpp:28-41): BVH build, then
 * emissive-triangle and environment alias tables.  opts may be NULL
 * (reference settings). */
pt_status pt_scene_build_accel(pt_scene* scene, const pt_bvh_options* opts);
/* SceneManager::uploadToDevice / DeviceScene::uploadFrom
 * (scene/device_scene.cu:8-57).  Resets frame_count to 0. */
pt_status pt_scene_upload(pt_scene* scene);

/* Setters of scene/scene_manager.h:32-36 (+ seed, samplePerFrame, camera as
 * SceneManager::updateCamera).  They take effect at the next upload/render. */
pt_status pt_scene_set_spp(pt_scene* scene, int spp);
pt_status pt_scene_set_sample_per_frame(pt_scene* scene, int spf);
pt_status pt_scene_set_max_depth(pt_scene* scene, int depth);
pt_status pt_scene_set_sampler(pt_scene* scene, int sampler);
pt_status pt_scene_set_seed(pt_scene* scene, uint32_t seed);
pt_status pt_scene_set_camera(pt_scene* scene, const pt_camera* cam);
pt_status pt_scene_get_camera(const pt_scene* scene, pt_camera* cam);
/* SceneManager::outputPath: the scene's "render.output" file (default
 * "output.hdr"), NUL-terminated into buf of cap bytes. */
pt_status pt_scene_output_path(const pt_scene* scene, char* buf, size_t cap);
/* buildCameraFrame (camera/camera.h:36-51) of the scene's current camera. */
pt_status pt_scene_camera_frame(const pt_scene* scene, pt_camera_frame* frame);

/* Read-only view of the scene's host arrays (HostScene fields), valid until
 * the scene is modified or destroyed.  textures are not included. */
typedef struct pt_host_scene_view {
    const pt_vertex* vertices; uint32_t num_vertices;
    const pt_face* faces; uint32_t num_faces;          /* original (load) order */
    const pt_material* materials; uint32_t num_materials;
    const pt_punctual_light* lights; uint32_t num_lights;
    const float* env_rgba; int32_t env_width, env_height;
    uint32_t num_textures;
    pt_camera camera;
    pt_render_settings settings;
} pt_host_scene_view;
pt_status pt_scene_host_view(const pt_scene* scene, pt_host_scene_view* out);
/* Texture i of the scene: RGBA floats (pointer valid like the view). */
pt_status pt_scene_texture(const pt_scene* scene, uint32_t i, pt_texture_desc* out);
/* Replaces the RGBA texels of texture i (same size) and of the environment
 * (i = 0xFFFFFFFF); takes effect at the next upload, light tables are rebuilt
 * by the next build_accel. */
pt_status pt_scene_set_texture_data(pt_scene* scene, uint32_t i, const float* rgba, size_t count);

/* Copies of the built acceleration data, for inspection and parity tests.
 * layout 0: the reference's depth-first order (left child = next node);
 * layout 1: the device order (siblings adjacent).  Pass NULL to query counts.
 * face_order[i] = original face index of BVH-ordered face i. */
pt_status pt_scene_get_bvh(const pt_scene* scene, int layout, float* nodes8, uint32_t* num_nodes,
                           uint32_t* face_order, uint32_t* num_faces);
/* 4-wide BVH collapsed from a binary one in the reference's node layout
 * (accel/bvh_common.h:16-24, as layout 0 above): 16 words per node
 * {P.xyz, S.x}, {S.y, S.z, qlo.x, qlo.y}, {qlo.z, qhi.x, qhi.y, qhi.z},
 * {child words}; child i's box on axis a is [fma(byte i of qlo.a, S.a, P.a),
 * fma(byte i of qhi.a, S.a, P.a)] and contains the child's exact box; child
 * words are a node index, 0x80000000 | first triangle for a leaf, or
 * 0xFFFFFFFF for an unused slot.  The binary nodes that become wide nodes
 * are the choice of least summed surface area (binary leaves are kept), so
 * the result is deterministic for a given input.  max_stack: traversal
 * stack entries a ray can need.  Host code (it has no reference counterpart:
 * the reference traverses the binary tree); the scene build runs the same
 * collapse on the GPU.  Pass NULL wide16 to query the count. */
pt_status pt_bvh_collapse_wide(const float* nodes8, uint32_t num_nodes, float* wide16, uint32_t* num_wide,
                               uint32_t* max_stack);
/* A copy of the scene's 4-wide BVH as pt_scene_build_accel made it on the
 * GPU (the layout above; equal, bit for bit, to pt_bvh_collapse_wide of the
 * layout-0 nodes), and its max_stack.  num_wide is 0 when the scene
 * traverses the binary tree only (non-finite bounds, or a tree deeper than
 * the traversal stack).  Pass NULL wide16 to query the count. */
pt_status pt_scene_get_wide_bvh(const pt_scene* scene, float* wide16, uint32_t* num_wide, uint32_t* max_stack);
/* Emissive-triangle and environment tables (lights/light_sampler.cpp:57-149).
 * Pass NULL to query counts.  alias entries are {float prob, uint32 alias}. */
pt_status pt_scene_get_light_tables(const pt_scene* scene,
                                    float* tri_pmf, uint32_t* n_tri_pmf,
                                    float* tri_alias, uint32_t* n_tri_alias,
                                    float* env_pmf, float* env_alias, uint32_t* n_env);

/* ------------------------------------------------------------------------ */
/* Film (film/film.cu): running-mean RGBA float, row 0 = bottom image row.   */
/* ------------------------------------------------------------------------ */
pt_status pt_film_create(int width, int height, pt_film** out);
void pt_film_destroy(pt_film* film);
/* Film::clear (film/film.cu:41-51) and Renderer::resetAccumulation
 * (renderer.cu:71-77) when a scene is given (its frame_count returns to 0). */
pt_status pt_film_reset(pt_film* film, pt_scene* scene);
/* Film::downloadAccumulated (film/film.cu:53-59); which 0 = acc, 1 = new. */
pt_status pt_film_download(const pt_film* film, int which, float* rgba_host);
/* Device pointer of the acc (which 0) or new (which 1) surface, W*H*4 floats. */
pt_status pt_film_device_ptr(const pt_film* film, int which, void** dev_ptr);
/* ImageIO::write (io/image_io.cpp:79-144): host RGBA floats, row 0 = bottom,
 * written with the rows flipped so the file's first row is the image top.
 * .hdr (Radiance RGBE) and .pfm keep floats; .png, .jpg/.jpeg (quality 90),
 * .bmp and .tga store clamp(pow(v, 1/2.2), 0, 1) * 255 + 0.5 per channel
 * (image_io.cpp:119-127).  Other extensions return PT_ERR_UNSUPPORTED. */
pt_status pt_image_write(const char* path, const float* rgba, int width, int height);
/* ImageIO::load (io/image_io.cpp:17-67): .hdr (RGBE) or .png/.jpg/.jpeg
 * (8-bit, decoded with pow(v/255, 2.2)); alpha = 1.  Call with rgba NULL for
 * the size, then with W*H*4 floats; row 0 = bottom (the film convention, so
 * a saved film reads back equal up to the format's precision). */
pt_status pt_image_read(const char* path, float* rgba, int* width, int* height);
/* The same as pt_image_read (kept for callers written against .hdr only). */
pt_status pt_image_read_hdr(const char* path, float* rgba, int* width, int* height);
/* Copies the acc (0) or new (1) surface into device memory dst (W*H*4 floats). */
pt_status pt_film_copy_device(const pt_film* film, int which, float* dst_device);

/* ------------------------------------------------------------------------ */
/* Rendering                                                                  */
/* ------------------------------------------------------------------------ */
/* Renderer::renderBatch (renderer.cu:20-45) -> Integrator::launch
 * (integrators/integrator.h:10-20, integrators/path.cu:244-255): renders
 * sample_per_frame samples per pixel into film, running-mean accumulates,
 * increments frame_count.  ms_out (may be NULL) gets the HIP-event time. */
pt_status pt_render_batch(pt_scene* scene, pt_film* film, int integrator, float* ms_out);
/* One batch of several scenes in one persistent wavefront schedule (no
 * reference counterpart).  Each job renders its scene's sample_per_frame
 * samples per pixel at the scene's frame_count into its own film (sizes may
 * differ) and advances that frame_count, with exactly the image
 * pt_render_batch(..., PT_INTEGRATOR_WAVEFRONT) would give.  The first job's
 * scene lends its wavefront buffers, configuration, stream and counters.
 * Films must be distinct; a scene may appear in several jobs. */
typedef struct pt_job {
    pt_scene* scene;
    pt_film* film;
} pt_job;
pt_status pt_render_jobs(const pt_job* jobs, int n, float* ms_out);
/* Renderer::renderAll (renderer.cu:87-111): batches until target_samples. */
pt_status pt_render_all(pt_scene* scene, pt_film* film, int integrator, float* total_ms);
/* Wavefront configuration of a scene (defaults: chunk_spp 64, blocks 64 and
 * 256, waves_per_simd 0, sort 1, graph 1, time_stages 0, shade_waves 3,
 * stack_lds 16, grabs 64 / 128 / 64, bvh_width 4). */
pt_status pt_scene_set_wavefront_config(pt_scene* scene, const pt_wavefront_config* cfg);
pt_status pt_scene_get_wavefront_config(const pt_scene* scene, pt_wavefront_config* cfg);
/* Autotuner: times up to max_trials configurations (coordinate descent over
 * chunk_spp, block sizes, sorting and grid size) on the uploaded scene with a
 * private film, keeps the fastest as the scene's configuration and returns it
 * with its GPU time per sample per pixel (ms).  The caller's films and
 * frame_count are untouched. */
pt_status pt_scene_tune_wavefront(pt_scene* scene, int max_trials, pt_wavefront_config* best, float* ms_per_spp);
pt_status pt_scene_wavefront_stats(const pt_scene* scene, pt_wavefront_stats* out);

/* ------------------------------------------------------------------------ */
/* Differentiation (no reference counterpart: the reference renders only)    */
/* ------------------------------------------------------------------------ */
/* Gradient buffers, device memory, accumulated with += (zero them first);
 * any may be NULL.  base_color: num_materials x 3 (the baseColor factor);
 * roughness, metallic: num_materials; textures: 4 floats per texel of the
 * texel pool (pt_scene_texture_offset), gradients of base-colour textures;
 * env: env_width x env_height x 4. */
typedef struct pt_adjoint_grads {
    float* base_color;
    float* roughness;
    float* metallic;
    float* textures;
    float* env;
} pt_adjoint_grads;
/* Path-replay backpropagation of batch `frame` (sample_per_frame samples per
 * pixel with the random numbers pt_render_batch used at that frame count):
 * image_grad (device, W*H*4 floats) is d loss / d (batch mean image, RGB;
 * the 4th channel is ignored).  replay_image (device W*H*4, may be NULL)
 * gets the batch image recomputed by the replay.  Sampling pdfs, MIS weights
 * and Russian roulette are detached; channels clamped at 100 carry no
 * gradient; FP32 atomics accumulate (last-bit nondeterminism). */
pt_status pt_render_adjoint(pt_scene* scene, uint32_t frame, const float* image_grad, const pt_adjoint_grads* grads,
                            float* replay_image, float* ms_out);
/* Form of the adjoint passes: PT_INTEGRATOR_WAVEFRONT (default: per chunk
 * of the persistent schedule a primal pass, then the replay with
 * material-sorted queues and shadow rays carrying their sensitivities) or
 * PT_INTEGRATOR_PATH (one thread per pixel, primal and replay in one kernel).
 * Both compute the same estimator. */
pt_status pt_scene_set_adjoint_integrator(pt_scene* scene, int integrator);
pt_status pt_scene_get_adjoint_integrator(const pt_scene* scene, int* integrator);
/* Many cameras of one scene (no reference counterpart).  pt_render_views
 * renders one batch per camera into films[i] (sized like cameras[i]) in one
 * wavefront schedule, with the random numbers of frames[i] (NULL: the
 * scene's frame_count for all; frame counts are not advanced).
 * pt_render_adjoint_views replays all views' batches in one persistent
 * schedule (one launch with PT_INTEGRATOR_PATH) and accumulates the
 * gradients of all of them into grads. */
pt_status pt_render_views(pt_scene* scene, int n, const pt_camera* cameras, pt_film* const* films,
                          const uint32_t* frames, float* ms_out);
pt_status pt_render_adjoint_views(pt_scene* scene, int n, const pt_camera* cameras, const uint32_t* frames,
                                  const float* const* image_grads, const pt_adjoint_grads* grads,
                                  float* const* replay_images, float* ms_out);
/* Keeping each path's radiance for the adjoint (no reference counterpart).
 * With keep on, every wavefront forward of this scene (pt_render_batch with
 * PT_INTEGRATOR_WAVEFRONT, pt_render_views) also stores each path's radiance
 * sum in device memory (16 bytes per path of the batch, freed when keep is
 * turned off); a wavefront adjoint (pt_render_adjoint, pt_render_adjoint_views)
 * of the same batch (the same cameras, frames and settings, and no update of
 * the scene's device data since) seeds its replay from it instead of running
 * the primal pass again.  Gradients and replay images are those of the
 * recomputing adjoint.  Off by default; without the memory for it a forward
 * renders without keeping.  reused (may be NULL): 1 when the last adjoint
 * seeded from kept radiance. */
pt_status pt_scene_set_keep_radiance(pt_scene* scene, int on);
pt_status pt_scene_get_keep_radiance(const pt_scene* scene, int* on, int* reused);
/* Sets frame_count (the sample-index offset of the next batch). */
pt_status pt_scene_set_frame(pt_scene* scene, uint32_t frame_count);
/* First texel and texel count of texture tex in the device texel pool. */
pt_status pt_scene_texture_offset(const pt_scene* scene, uint32_t tex, uint64_t* texel_offset,
                                  uint64_t* texel_count);
/* Parameter updates of an uploaded scene without a re-upload: material
 * baseColor (num_materials x 3), roughness, metallic from host arrays (NULL
 * keeps the field); texels of texture tex (0xFFFFFFFF = the environment)
 * from host or device memory (on_device), RGBA floats of the same size.
 * The environment's sampling tables keep their old distribution until
 * pt_scene_rebuild_env_tables. */
pt_status pt_scene_update_materials(pt_scene* scene, const float* base_color, const float* roughness,
                                    const float* metallic);
pt_status pt_scene_update_texture(pt_scene* scene, uint32_t tex, const float* rgba, size_t count, int on_device);
pt_status pt_scene_rebuild_env_tables(pt_scene* scene);
/* The same tables built on the GPU from the device environment texels
 * (light_sampler.cpp:112-149 weights, no host round trip): the pmf equals the
 * host build's up to rounding and the alias table samples the same
 * distribution, but its layout differs from the reference's Vose table, so a
 * given random number picks another texel (the estimator is unchanged).
 * For inverse rendering, where the environment changes every iteration. */
pt_status pt_scene_build_env_tables_device(pt_scene* scene);
/* Emission updates of an uploaded scene: material emissive color
 * (num_materials x 3) and strength (num_materials) from host arrays (NULL
 * keeps the field).  The emissive-triangle sampling tables keep their old
 * distribution until one of the two rebuilds below. */
pt_status pt_scene_update_emission(pt_scene* scene, const float* emissive, const float* strength);
/* Rebuild of the emissive-triangle tables on the host, bit-identical to the
 * reference's (lights/light_sampler.cpp:57-110, called from
 * scene/scene_manager.cpp:31-34 at load), and uploaded. */
pt_status pt_scene_rebuild_emissive_tables(pt_scene* scene);
/* The same tables built on the GPU from the device vertices, BVH-ordered
 * faces and materials (no host round trip): the emissive list (face indices
 * in face order) equals the host build's, the per-face pmf equals it up to
 * rounding, and the alias table samples the same distribution in another
 * layout.  The number of emissive faces may change (an emission set to 0
 * drops its faces).  For inverse rendering of emission. */
pt_status pt_scene_build_emissive_tables_device(pt_scene* scene);
/* Alias table of n >= 1 non-negative device weights (not all zero), built on
 * the GPU: pmf[i] = w[i] / sum (n floats), alias = n {prob, index bits}
 * pairs (2n floats), all device memory.  Synchronous. */
pt_status pt_alias_build_device(const float* weights, uint32_t n, float* pmf, float* alias);
/* pt_alias_build_device without allocating: its working buffers (40 bytes
 * per weight) sit in a device array of the library, one per GPU, held until
 * the process ends, so n <= PT_ALIAS_NOALLOC_MAX_WEIGHTS (more is
 * PT_ERR_INVALID_ARG).  No hipMalloc/hipFree: for callers that cannot
 * allocate (tight loops that rebuild an environment's tables after each
 * update, harnesses that forbid allocation).  Same tables as
 * pt_alias_build_device, bit for bit.  Calls on one GPU share the array:
 * calls that overlap on one GPU race.  Synchronous.  No counterpart in the
 * reference, whose tables are built on the host. */
#define PT_ALIAS_NOALLOC_MAX_WEIGHTS (1u << 20)
pt_status pt_alias_build_device_noalloc(const float* weights, uint32_t n, float* pmf, float* alias);
/* Counter control: when enabled, render kernels count rays, node visits and
 * triangle tests (slower; used to price algorithmic bytes). */
pt_status pt_counters_enable(pt_scene* scene, int enable);
pt_status pt_counters_get(pt_scene* scene, pt_counters* out);
pt_status pt_counters_reset(pt_scene* scene);

/* ------------------------------------------------------------------------ */
/* Probes of single hot-path stages on the device                             */
/* ------------------------------------------------------------------------ */
/* Sampler::get1D (samplers/sampler.cuh:25-29) for n (x, y, sample_index)
 * triples over dims [0, num_dims); out_u32 gets the raw 32-bit value before
 * the float conversion, out_f the float.  sampler 0 = PCG (sample_index is
 * the frame/sample pair encoded as frame*65536+sample), 1 = Sobol. */
pt_status pt_probe_sampler(int sampler, uint32_t seed, const int32_t* xy_idx, int n, int num_dims,
                           uint32_t* out_u32, float* out_f);
/* Closest-hit (intersectScene, accel/bvh_traversal.cuh:43-96) or any-hit
 * (occludedScene, :98-133) for n rays {ox,oy,oz,dx,dy,dz,tmax,_} given as 8
 * floats each, host buffers.  Directions are normalised as Ray's constructor
 * does (core/ray.h:13).  t_out gets the hit distance (FLT_MAX on a miss; any-
 * hit: 1 occluded / 0 not), face_out the ORIGINAL face index or -1,
 * bary_out (may be NULL) 2 floats per ray.  any_hit bit 0: any-hit; bit 1:
 * traverse the scene's 4-wide BVH instead of the binary one. */
/* Disney BSDF (disney_bsdf.cuh:53-122 fetch + clamps without textures,
 * :209-321 eval, :325-410 sample) of n materials on the device: dirs9 holds
 * wo, wi, n per case, u3 the three uniforms a sample consumes (lobe choice,
 * then the warp pair).  out12 per case: f.xyz, pdf, sampled wi.xyz,
 * weight.xyz, pdf, flags (1 valid, 2 delta). */
pt_status pt_probe_bsdf(const pt_material* mats, int n, const float* dirs9, const int32_t* entering, const float* u3,
                        float* out12);
pt_status pt_trace_rays(pt_scene* scene, const float* rays8, int n, int any_hit,
                        float* t_out, int32_t* face_out, float* bary_out);

/* Streams `bytes` (rounded down to 16 B) from one device buffer to another
 * `iters` times with a plain vectorised copy kernel and reports GB/s (read +
 * write) from HIP events.  No reference counterpart: it calibrates HBM
 * traffic counters and states the achievable bandwidth of the box. */
pt_status pt_probe_copy(size_t bytes, int iters, float* gbps);

/* Compares the triangle test's reciprocal (hardware reciprocal plus one
 * Newton step) with IEEE division, bit for bit, for every float of either
 * sign with magnitude in [2^-126, 2^126); returns the number of differing
 * results and the number compared (2 x 252 x 2^23).  No reference
 * counterpart: it checks that the triangle test's 1/det equals the
 * reference's 1.0f / det (geometry/triangle.h:27). */
pt_status pt_probe_reciprocal(uint64_t* mismatches, uint64_t* compared);

#ifdef __cplusplus
}
#endif
#endif /* PT_AMD_H */

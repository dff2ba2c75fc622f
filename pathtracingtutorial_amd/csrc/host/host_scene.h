// Host-side scene model and the host stages of the accel build.
//
// HostScene mirrors the reference's HostScene (Part_2/src/scene/scene.h:24-35):
// vertex/face/material arrays in the reference byte layouts (declared in
// include/pt_amd.h), textures as RGBA float images, punctual lights, camera,
// environment image and render settings.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "pt_amd.h"

namespace ptamd {

struct Image {
    int width = 0, height = 0;
    int wrap_s = 10497, wrap_t = 10497;  // GL enums, io/image_io.h:14-15
    int linear = 0;                      // 1: skip the pow(2.2) colour decode
    std::vector<float> rgba;             // width*height*4, row 0 first
    bool empty() const { return width <= 0 || height <= 0 || rgba.empty(); }
};

struct HostScene {
    std::vector<pt_vertex> vertices;
    std::vector<pt_face> faces;
    std::vector<pt_material> materials;
    std::vector<Image> textures;
    std::vector<pt_punctual_light> lights;
    pt_camera camera{};
    Image env;
    pt_render_settings settings{};
    std::string output_file = "output.hdr";
    std::string path;
    // BVH options carried by generated workloads (bins/leaf sizes of the config).
    pt_bvh_options bvh_opts{};
    bool has_bvh_opts = false;
};

// Defaults of the reference structs.
pt_material default_material();               // materials/material.h:8-81
pt_camera default_camera();                   // camera/camera.h:12-21
pt_punctual_light default_light();            // lights/punctual_light.h:14-25
pt_render_settings default_settings();        // scene/scene.h:16-22
void init_scene_defaults(HostScene& s);

// Scene loading (scene/scene_loader.cpp).  Return false with *err set.
bool load_scene_json(const std::string& path, HostScene& scene, std::string* err);
bool load_scene(const std::string& path, HostScene& scene, std::string* err);  // .json, .gltf, .glb
bool load_obj(const std::string& path, int material_id, const float xform[16],
              std::vector<pt_vertex>& verts, std::vector<pt_face>& faces, std::string* err);
bool load_hdr(const std::string& path, Image& img, std::string* err);
bool write_hdr(const std::string& path, int w, int h, const float* rgba, bool flip_y, std::string* err);
bool write_pfm(const std::string& path, int w, int h, const float* rgba, std::string* err);
// ImageIO::load / ImageIO::write (io/image_io.cpp): every format the reference reads and writes.
bool load_image(const std::string& path, Image& img, std::string* err);
bool write_image(const std::string& path, int w, int h, const float* rgba, bool flip_y, std::string* err);
uint8_t ldr_byte(float v);
void build_quad(const float a[3], const float b[3], const float c[3], const float d[3], int mat,
                std::vector<pt_vertex>& verts, std::vector<pt_face>& faces);
void auto_fit_camera(const std::vector<pt_vertex>& verts, pt_camera& cam);

// 4x4 row-major transforms (math/mat4.h semantics: T * R(axis, deg) * S).
void mat4_identity(float m[16]);
void mat4_mul(const float a[16], const float b[16], float out[16]);
void mat4_trs(const float t[3], const float axis[3], float angle_deg, float scale, float out[16]);
bool mat4_inverse(const float m[16], float out[16]);

// Procedural workloads (seeded).
bool generate_cornell(HostScene& s);
bool generate_diff_cornell(HostScene& s, uint64_t seed);
bool generate_batch_job(HostScene& s, uint64_t seed);
bool generate_diff_views(HostScene& s, uint64_t seed);
bool generate_spheres(HostScene& s, uint64_t seed, int instances, int level, int env_w, int env_h);
void generate_sky_env(Image& env, int w, int h, uint64_t seed);

// Camera frame (camera/camera.h:36-51).
pt_camera_frame camera_frame(const pt_camera& cam);

// Light tables (lights/light_sampler.cpp).  faces are in BVH order.
struct AliasEntry { float prob; uint32_t alias; };
struct EmissiveRef { uint32_t tri; float pmf; float cdf; float pad; };
struct LightTables {
    std::vector<EmissiveRef> emissive;
    std::vector<float> tri_pmf;          // per face
    std::vector<AliasEntry> tri_alias;
    std::vector<float> env_pmf;
    std::vector<AliasEntry> env_alias;
    int env_w = 0, env_h = 0;
};
std::vector<AliasEntry> build_alias_table(const std::vector<float>& weights);
void build_emissive_tables(const std::vector<pt_vertex>& v, const std::vector<pt_face>& f,
                           const std::vector<pt_material>& m, LightTables& out);
void build_env_tables(const Image& env, LightTables& out);

bool material_is_emissive(const pt_material& m);

}  // namespace ptamd

// Driver over the reference renderer's own host-compilable sources, compiled
// where they lie (see oracle/Makefile) into oracle/_ref/libref.so.  Used only
// by tests/golden/make_golden.py to write golden vectors; never shipped.
//
// Every function calls the reference code; the only glue is argument
// marshalling.  Sources used: accel/bvh_builder.cpp, accel/bvh_traversal.cuh,
// geometry/triangle.h, samplers/sobol.cuh, lights/light_sampler.cpp,
// lights/envmap.h, materials/microfacet.cuh, math/sampling.h, math/frame.h,
// io/mesh_loader.cpp (+ vendored tinyobjloader), io/image_io.cpp (+ vendored
// stb), geometry/shapes.cpp, math/mat4.h.
#include <cstddef>
#include <cstring>
#include <string>
#include <vector>

#include "accel/bvh_builder.h"
#include "accel/bvh_traversal.cuh"
#include "geometry/shapes.h"
#include "io/image_io.h"
#include "io/mesh_loader.h"
#include "lights/envmap.h"
#include "lights/light_sampler.h"
#include "lights/punctual_light.h"
#include "materials/material.h"
#include "materials/microfacet.cuh"
#include "math/frame.h"
#include "math/mat4.h"
#include "math/sampling.h"
#include "samplers/sobol.cuh"

using namespace pt;

namespace {
struct RefScene {
    std::vector<Vertex> v;
    std::vector<TriangleFace> f;
    BVH bvh;
};
}  // namespace

extern "C" {

// ---- layout of the data contract ---------------------------------------------
void ref_layout(uint32_t* out) {
    out[0] = sizeof(Vertex);
    out[1] = sizeof(TriangleFace);
    out[2] = sizeof(Material);
    out[3] = sizeof(PunctualLight);
    out[4] = sizeof(BVHNode);
    out[5] = sizeof(AliasEntry);
    out[6] = sizeof(EmissiveTriangleRef);
    out[7] = offsetof(Vertex, normal);
    out[8] = offsetof(Vertex, uv);
    out[9] = offsetof(Vertex, tangent);
    out[10] = offsetof(Material, metallic);
    out[11] = offsetof(Material, emissive);
    out[12] = offsetof(Material, emissiveStrength);
    out[13] = offsetof(Material, alphaMode);
    out[14] = offsetof(Material, doubleSided);
    out[15] = offsetof(Material, attColorB);
    out[16] = offsetof(Material, uvRotation);
    out[17] = offsetof(Material, baseColorTexId);
    out[18] = offsetof(Material, sheenRoughnessTexId);
    out[19] = offsetof(PunctualLight, color);
    out[20] = offsetof(PunctualLight, direction);
    out[21] = offsetof(PunctualLight, type);
    out[22] = offsetof(Material, specTrans);
    out[23] = offsetof(Material, thicknessFactor);
}

// ---- default Material (materials/material.h) --------------------------------
void ref_default_material(void* out) {
    Material m;
    std::memcpy(out, &m, sizeof(m));
}

// ---- Sobol (samplers/sobol.cuh) ----------------------------------------------
float ref_sobol_get1d(int px, int py, uint32_t index, uint32_t dim) {
    SobolSampler s;
    s.start(px, py, index);
    float v = 0.0f;
    for (uint32_t d = 0; d <= dim; ++d) v = s.get1D();
    return v;
}
uint32_t ref_sobol_hashmix(uint32_t x) { return hashMix(x); }
uint32_t ref_sobol_owen(uint32_t v, uint32_t seed) { return owenScramble(v, seed); }
uint32_t ref_sobol_dimseed(int px, int py, uint32_t dim) { return dimensionSeed(px, py, dim); }
uint32_t ref_reverse_bits(uint32_t x) { return reverseBits32(x); }

// ---- BVH build + traversal ---------------------------------------------------
void* ref_scene_build(const void* verts64, uint32_t nv, const void* faces16, uint32_t nf) {
    RefScene* s = new RefScene();
    s->v.resize(nv);
    std::memcpy(s->v.data(), verts64, sizeof(Vertex) * nv);
    s->f.resize(nf);
    std::memcpy(s->f.data(), faces16, sizeof(TriangleFace) * nf);
    // the original face index rides in materialID through the in-place reorder
    for (uint32_t i = 0; i < nf; ++i) s->f[i].materialID = static_cast<int>(i);
    BVHBuilder b;
    s->bvh = b.build(s->v, s->f, 0, nf);
    return s;
}
void ref_scene_free(void* p) { delete static_cast<RefScene*>(p); }
uint32_t ref_bvh_num_nodes(void* p) { return static_cast<uint32_t>(static_cast<RefScene*>(p)->bvh.nodes.size()); }
void ref_bvh_get(void* p, void* nodes32, uint32_t* face_order) {
    RefScene* s = static_cast<RefScene*>(p);
    std::memcpy(nodes32, s->bvh.nodes.data(), s->bvh.nodes.size() * sizeof(BVHNode));
    for (size_t i = 0; i < s->f.size(); ++i) face_order[i] = static_cast<uint32_t>(s->f[i].materialID);
}
void ref_trace(void* p, const float* rays8, int n, int any, float* t, int32_t* face, float* bary) {
    RefScene* s = static_cast<RefScene*>(p);
    for (int i = 0; i < n; ++i) {
        const float* r = rays8 + 8 * static_cast<size_t>(i);
        Ray ray(vec3(r[0], r[1], r[2]), vec3(r[3], r[4], r[5]));
        if (any) {
            t[i] = occludedScene(ray, s->bvh.nodes.data(), s->v.data(), s->f.data(), r[6]) ? 1.0f : 0.0f;
            face[i] = -1;
            continue;
        }
        SurfaceInteraction si = intersectScene(ray, s->bvh.nodes.data(), s->v.data(), s->f.data());
        t[i] = si.distance;
        face[i] = si.hit ? si.materialId : -1;  // materialID carries the original face index
        if (bary) { bary[2 * i] = si.baryU; bary[2 * i + 1] = si.baryV; }
    }
}

// ---- light tables (lights/light_sampler.cpp) ---------------------------------
uint32_t ref_env_sampler(const float* rgba, int w, int h, float* pmf, float* alias8) {
    Image img;
    img.width = w;
    img.height = h;
    img.channel = 4;
    img.data.assign(rgba, rgba + static_cast<size_t>(w) * h * 4);
    HostLightSampler ls;
    buildEnvironmentSampler(img, ls);
    if (pmf) std::memcpy(pmf, ls.envPmf.data(), ls.envPmf.size() * 4);
    if (alias8) std::memcpy(alias8, ls.envAlias.data(), ls.envAlias.size() * 8);
    return static_cast<uint32_t>(ls.envPmf.size());
}
uint32_t ref_emissive_sampler(const void* verts64, uint32_t nv, const void* faces16, uint32_t nf, const void* mats,
                              uint32_t nm, float* tri_pmf, float* alias8, uint32_t* tri_index) {
    std::vector<Vertex> v(nv);
    std::memcpy(v.data(), verts64, sizeof(Vertex) * nv);
    std::vector<TriangleFace> f(nf);
    std::memcpy(f.data(), faces16, sizeof(TriangleFace) * nf);
    std::vector<Material> m(nm);
    std::memcpy(m.data(), mats, sizeof(Material) * nm);
    HostLightSampler ls;
    buildEmissiveTriangleSampler(v, f, m, ls);
    if (tri_pmf) std::memcpy(tri_pmf, ls.triangleSelectionPmf.data(), ls.triangleSelectionPmf.size() * 4);
    if (alias8) std::memcpy(alias8, ls.triangleAlias.data(), ls.triangleAlias.size() * 8);
    if (tri_index)
        for (size_t i = 0; i < ls.emissiveTriangles.size(); ++i) tri_index[i] = ls.emissiveTriangles[i].triangleIndex;
    return static_cast<uint32_t>(ls.emissiveTriangles.size());
}

// ---- env mapping (lights/envmap.h) --------------------------------------------
void ref_dir_to_uv(const float* d, float* uv) {
    vec2 r = sampleSphericalMap(vec3(d[0], d[1], d[2]));
    uv[0] = r.x;
    uv[1] = r.y;
}
void ref_uv_to_dir(const float* uv, float* d) {
    vec3 r = directionFromSphericalMap(vec2(uv[0], uv[1]));
    d[0] = r.x; d[1] = r.y; d[2] = r.z;
}

// ---- microfacet / sampling / frame ---------------------------------------------
float ref_d_ggx(float noh, float a) { return D_GGX(noh, a); }
float ref_d_ggx_aniso(float noh, float hx, float hy, float ax, float ay) { return D_GGX_Aniso(noh, hx, hy, ax, ay); }
float ref_g1(float nov, float a) { return G1_Smith(nov, a); }
float ref_g2(float nol, float nov, float a) { return G2_SmithCorrelated(nol, nov, a); }
float ref_f_schlick(float c, float f0) { return F_Schlick(c, f0); }
void ref_f_schlick3(float c, const float* f0, float* out) {
    vec3 r = F_Schlick(c, vec3(f0[0], f0[1], f0[2]));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
float ref_f_dielectric(float c, float eta) { return F_Dielectric(c, eta); }
void ref_sample_vndf(const float* wo, float ax, float ay, float u1, float u2, float* h) {
    vec3 r = sampleVNDF(vec3(wo[0], wo[1], wo[2]), ax, ay, u1, u2);
    h[0] = r.x; h[1] = r.y; h[2] = r.z;
}
float ref_pdf_vndf(const float* wo, const float* h, float ax, float ay) {
    return pdfVNDF(vec3(wo[0], wo[1], wo[2]), vec3(h[0], h[1], h[2]), ax, ay);
}
int ref_refract(const float* wi, const float* n, float eta, float* wt) {
    vec3 r;
    bool ok = refractDir(vec3(wi[0], wi[1], wi[2]), vec3(n[0], n[1], n[2]), eta, r);
    wt[0] = r.x; wt[1] = r.y; wt[2] = r.z;
    return ok ? 1 : 0;
}
void ref_cosine_hemisphere(float u1, float u2, float* d) {
    vec3 r = sampleCosineHemisphere(u1, u2);
    d[0] = r.x; d[1] = r.y; d[2] = r.z;
}
float ref_power_heuristic(float a, float b) { return powerHeuristic(a, b); }
void ref_frame(const float* n, const float* v, float* local, float* world) {
    Frame fr(vec3(n[0], n[1], n[2]));
    vec3 l = fr.toLocal(vec3(v[0], v[1], v[2]));
    vec3 w = fr.toWorld(vec3(v[0], v[1], v[2]));
    local[0] = l.x; local[1] = l.y; local[2] = l.z;
    world[0] = w.x; world[1] = w.y; world[2] = w.z;
}

// ---- loaders --------------------------------------------------------------------
// MeshLoader::loadObj with T*R*S built by mat4 as scene_loader.cpp:70-72 does.
int ref_load_obj(const char* path, int mat, const float* t3, const float* axis3, float angle_deg, float scale,
                 void* verts64_out, uint32_t cap_v, uint32_t* nv, void* faces16_out, uint32_t cap_f, uint32_t* nf) {
    mat4 x = mat4::translation(vec3(t3[0], t3[1], t3[2])) *
             mat4::rotation(angle_deg * Pi / 180.0f, vec3(axis3[0], axis3[1], axis3[2])) *
             mat4::scaling(scale, scale, scale);
    std::vector<Vertex> v;
    std::vector<TriangleFace> f;
    bool ok = MeshLoader::loadObj(path, mat, x, v, f);
    *nv = static_cast<uint32_t>(v.size());
    *nf = static_cast<uint32_t>(f.size());
    if (verts64_out && v.size() <= cap_v) std::memcpy(verts64_out, v.data(), v.size() * sizeof(Vertex));
    if (faces16_out && f.size() <= cap_f) std::memcpy(faces16_out, f.data(), f.size() * sizeof(TriangleFace));
    return ok ? 1 : 0;
}
// ImageIO::load (stb_image) of an .hdr: returns w*h*4 floats into out when it fits.
int ref_load_image(const char* path, float* out, size_t cap, int* w, int* h) {
    Image img;
    if (!ImageIO::load(img, path)) return 0;
    *w = img.width;
    *h = img.height;
    if (out && img.data.size() <= cap) std::memcpy(out, img.data.data(), img.data.size() * 4);
    return 1;
}
void ref_build_quad(const float* a, const float* b, const float* c, const float* d, int mat, void* verts64,
                    void* faces16) {
    std::vector<Vertex> v;
    std::vector<TriangleFace> f;
    shapes::buildQuad(vec3(a[0], a[1], a[2]), vec3(b[0], b[1], b[2]), vec3(c[0], c[1], c[2]),
                      vec3(d[0], d[1], d[2]), mat, v, f);
    std::memcpy(verts64, v.data(), v.size() * sizeof(Vertex));
    std::memcpy(faces16, f.data(), f.size() * sizeof(TriangleFace));
}

}  // extern "C"

// Scene description loading: JSON schema, OBJ meshes, Radiance .hdr images,
// transforms, quads, camera frames.
//
// Behaviour follows the reference loader (Part_2/src/scene/scene_loader.cpp,
// io/mesh_loader.cpp, io/image_io.cpp, geometry/shapes.cpp, math/mat4.h,
// camera/camera.h); each function cites the lines it mirrors.
#include <cerrno>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <unordered_map>

#include "gltf.h"
#include "host_scene.h"
#include "image_codec.h"
#include "json.h"

namespace ptamd {

// ---------------------------------------------------------------- defaults --

pt_material default_material() {  // materials/material.h:8-81
    pt_material m;
    std::memset(&m, 0, sizeof(m));
    m.base_color = {0.8f, 0.8f, 0.8f, 0.0f};
    m.metallic = 0.0f;
    m.roughness = 0.5f;
    m.specular_tint = 0.0f;
    m.anisotropic = 0.0f;
    m.sheen = 0.0f;
    m.sheen_tint = 0.5f;
    m.clearcoat = 0.0f;
    m.clearcoat_roughness = 0.03f;
    m.spec_trans = 0.0f;
    m.ior = 1.5f;
    m.subsurface = 0.0f;
    m.specular_factor = 1.0f;
    m.specular_color_r = m.specular_color_g = m.specular_color_b = 1.0f;
    m.anisotropy_rotation = 0.0f;
    m.normal_scale = 1.0f;
    m.emissive = {0.0f, 0.0f, 0.0f, 0.0f};
    m.emissive_strength = 1.0f;
    m.alpha_cutoff = 0.0f;
    m.alpha_mode = 0;
    m.base_alpha = 1.0f;
    m.double_sided = 0;
    m.sheen_color_r = m.sheen_color_g = m.sheen_color_b = 1.0f;
    m.occlusion_strength = 1.0f;
    m.thickness_factor = 0.0f;
    m.attenuation_distance = 1e30f;
    m.att_color_r = m.att_color_g = m.att_color_b = 1.0f;
    m.uv_offset_x = m.uv_offset_y = 0.0f;
    m.uv_scale_x = m.uv_scale_y = 1.0f;
    m.uv_rotation = 0.0f;
    m.base_color_tex_id = m.metallic_roughness_tex_id = m.normal_tex_id = m.emissive_tex_id = -1;
    m.occlusion_tex_id = m.clearcoat_tex_id = m.clearcoat_roughness_tex_id = m.clearcoat_normal_tex_id = -1;
    m.transmission_tex_id = m.sheen_color_tex_id = m.sheen_roughness_tex_id = -1;
    m._pad3 = 0;
    return m;
}

pt_camera default_camera() {  // camera/camera.h:12-21
    pt_camera c;
    std::memset(&c, 0, sizeof(c));
    c.position = {0.0f, 0.0f, 4.0f, 0.0f};
    c.look_at = {0.0f, 0.0f, 0.0f, 0.0f};
    c.up = {0.0f, 1.0f, 0.0f, 0.0f};
    c.fov = 90.0f;
    c.aperture = 0.0f;
    c.focal_distance = 10.0f;
    c.width = 1600;
    c.height = 1600;
    return c;
}

pt_punctual_light default_light() {  // lights/punctual_light.h:14-25
    pt_punctual_light l;
    std::memset(&l, 0, sizeof(l));
    l.range = FLT_MAX;
    l.color = {1.0f, 1.0f, 1.0f, 0.0f};
    l.intensity = 1.0f;
    l.direction = {0.0f, -1.0f, 0.0f, 0.0f};
    l.inner_cone = 0.0f;
    l.outer_cone = 0.7071068f;
    l.type = 0;
    return l;
}

pt_render_settings default_settings() {  // scene/scene.h:16-22
    pt_render_settings s;
    std::memset(&s, 0, sizeof(s));
    s.sample_per_frame = 4;
    s.target_samples = 100000;
    s.max_depth = 8;
    s.sampler = 1;
    s.seed = 0;
    return s;
}

void init_scene_defaults(HostScene& s) {
    s.camera = default_camera();
    s.settings = default_settings();
    s.output_file = "output.hdr";
}

bool material_is_emissive(const pt_material& m) {  // materials/material.h:83-86
    return m.emissive_strength > 0.0f && (m.emissive.x > 0.0f || m.emissive.y > 0.0f || m.emissive.z > 0.0f);
}

// ------------------------------------------------------------ small vector --

namespace {
struct V3 {
    float x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 mul(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float length(V3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
// vecmath.h:96: divide by the length, zero vector stays zero.
inline V3 normalize(V3 a) {
    float l = length(a);
    return l > 0 ? V3{a.x / l, a.y / l, a.z / l} : V3{0.0f, 0.0f, 0.0f};
}
inline V3 from(const pt_vec3p& p) { return {p.x, p.y, p.z}; }
inline pt_vec3p to3p(V3 v) { return {v.x, v.y, v.z, 0.0f}; }
}  // namespace

// ------------------------------------------------------------------- mat4 --

void mat4_identity(float m[16]) {
    for (int i = 0; i < 16; ++i) m[i] = (i % 5 == 0) ? 1.0f : 0.0f;
}

void mat4_mul(const float a[16], const float b[16], float out[16]) {
    float r[16];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            float s = 0.0f;
            for (int k = 0; k < 4; ++k) s += a[i * 4 + k] * b[k * 4 + j];
            r[i * 4 + j] = s;
        }
    std::memcpy(out, r, sizeof(r));
}

// Axis-angle rotation (Rodrigues), translation and uniform scale composed as
// T * R * S, the order of scene_loader.cpp:70-72 (math/mat4.h:139-163).
void mat4_trs(const float t[3], const float axis[3], float angle_deg, float scale, float out[16]) {
    float T[16], R[16], S[16], TR[16];
    mat4_identity(T);
    T[3] = t[0]; T[7] = t[1]; T[11] = t[2];
    mat4_identity(S);
    S[0] = scale; S[5] = scale; S[10] = scale;
    mat4_identity(R);
    const float pi = 3.14159265358979323846f;
    float rad = angle_deg * pi / 180.0f;
    float len = std::sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
    if (len >= 1e-8f) {
        float x = axis[0] / len, y = axis[1] / len, z = axis[2] / len;
        float c = cosf(rad), s = sinf(rad), omc = 1.0f - c;
        R[0] = c + omc * x * x;      R[1] = omc * x * y - s * z;  R[2] = omc * x * z + s * y;  R[3] = 0;
        R[4] = omc * y * x + s * z;  R[5] = c + omc * y * y;      R[6] = omc * y * z - s * x;  R[7] = 0;
        R[8] = omc * z * x - s * y;  R[9] = omc * z * y + s * x;  R[10] = c + omc * z * z;     R[11] = 0;
        R[12] = 0; R[13] = 0; R[14] = 0; R[15] = 1;
    }
    mat4_mul(T, R, TR);
    mat4_mul(TR, S, out);
}

// Gauss-Jordan elimination with partial pivoting; singular -> identity, as
// math/mat4.h:81-108 returns.
bool mat4_inverse(const float m[16], float out[16]) {
    float A[16], inv[16];
    std::memcpy(A, m, sizeof(A));
    mat4_identity(inv);
    for (int i = 0; i < 4; ++i) {
        int piv = i;
        float best = std::fabs(A[i * 4 + i]);
        for (int j = i + 1; j < 4; ++j) {
            float c = std::fabs(A[j * 4 + i]);
            if (c > best) { best = c; piv = j; }
        }
        if (best < 1e-8f) { mat4_identity(out); return false; }
        if (piv != i)
            for (int k = 0; k < 4; ++k) {
                std::swap(A[i * 4 + k], A[piv * 4 + k]);
                std::swap(inv[i * 4 + k], inv[piv * 4 + k]);
            }
        float p = A[i * 4 + i];
        for (int k = 0; k < 4; ++k) { A[i * 4 + k] /= p; inv[i * 4 + k] /= p; }
        for (int j = 0; j < 4; ++j) {
            if (j == i) continue;
            float f = A[j * 4 + i];
            for (int k = 0; k < 4; ++k) { A[j * 4 + k] -= f * A[i * 4 + k]; inv[j * 4 + k] -= f * inv[i * 4 + k]; }
        }
    }
    std::memcpy(out, inv, sizeof(inv));
    return true;
}

static V3 xform_point(const float m[16], V3 v) {
    return {m[0] * v.x + m[1] * v.y + m[2] * v.z + m[3], m[4] * v.x + m[5] * v.y + m[6] * v.z + m[7],
            m[8] * v.x + m[9] * v.y + m[10] * v.z + m[11]};
}
static V3 xform_vector(const float m[16], V3 v) {
    return {m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
            m[8] * v.x + m[9] * v.y + m[10] * v.z};
}

// ------------------------------------------------------------------ shapes --

// geometry/shapes.cpp:5-22: two triangles (a,b,c) and (a,c,d), flat normal
// cross(b-a, c-a), tangent along b-a, uv on the unit square.
void build_quad(const float a[3], const float b[3], const float c[3], const float d[3], int mat,
                std::vector<pt_vertex>& verts, std::vector<pt_face>& faces) {
    V3 A{a[0], a[1], a[2]}, B{b[0], b[1], b[2]}, C{c[0], c[1], c[2]}, D{d[0], d[1], d[2]};
    V3 n = normalize(cross(sub(B, A), sub(C, A)));
    V3 t = normalize(sub(B, A));
    uint32_t base = static_cast<uint32_t>(verts.size());
    const V3 P[4] = {A, B, C, D};
    const float U[4] = {0.0f, 1.0f, 1.0f, 0.0f}, Vv[4] = {0.0f, 0.0f, 1.0f, 1.0f};
    for (int i = 0; i < 4; ++i) {
        pt_vertex v;
        std::memset(&v, 0, sizeof(v));
        v.position = to3p(P[i]);
        v.normal = to3p(n);
        v.u = U[i];
        v.v = Vv[i];
        v.tangent[0] = t.x; v.tangent[1] = t.y; v.tangent[2] = t.z; v.tangent[3] = 1.0f;
        verts.push_back(v);
    }
    faces.push_back({base + 0, base + 1, base + 2, mat});
    faces.push_back({base + 0, base + 2, base + 3, mat});
}

// scene_loader.cpp:212-238.
void auto_fit_camera(const std::vector<pt_vertex>& verts, pt_camera& cam) {
    if (verts.empty()) return;
    V3 lo{1e30f, 1e30f, 1e30f}, hi{-1e30f, -1e30f, -1e30f};
    for (const auto& v : verts) {
        lo.x = std::fmin(lo.x, v.position.x); lo.y = std::fmin(lo.y, v.position.y); lo.z = std::fmin(lo.z, v.position.z);
        hi.x = std::fmax(hi.x, v.position.x); hi.y = std::fmax(hi.y, v.position.y); hi.z = std::fmax(hi.z, v.position.z);
    }
    V3 center = mul(add(lo, hi), 0.5f);
    float radius = length(sub(hi, lo)) * 0.5f;
    if (radius < 1e-5f) radius = 1.0f;
    const float pi = 3.14159265358979323846f;
    float fov = cam.fov * pi / 180.0f;
    float dist = radius / std::tan(fov * 0.5f) * 1.2f;
    cam.position = to3p(add(center, V3{0.0f, radius * 0.3f, dist}));
    cam.look_at = to3p(center);
    cam.up = {0.0f, 1.0f, 0.0f, 0.0f};
}

// camera/camera.h:36-51.
pt_camera_frame camera_frame(const pt_camera& cam) {
    V3 fwd = normalize(sub(from(cam.look_at), from(cam.position)));
    V3 rgt = normalize(cross(fwd, from(cam.up)));
    V3 cup = cross(rgt, fwd);
    pt_camera_frame f;
    std::memset(&f, 0, sizeof(f));
    f.position[0] = cam.position.x; f.position[1] = cam.position.y; f.position[2] = cam.position.z;
    f.forward[0] = fwd.x; f.forward[1] = fwd.y; f.forward[2] = fwd.z;
    f.right[0] = rgt.x; f.right[1] = rgt.y; f.right[2] = rgt.z;
    f.up[0] = cup.x; f.up[1] = cup.y; f.up[2] = cup.z;
    const float pi = 3.14159265358979323846f;
    f.tan_half_fov = std::tan(cam.fov * pi / 360.0f);
    f.aspect = static_cast<float>(cam.width) / cam.height;
    f.aperture = cam.aperture;
    f.focal_distance = cam.focal_distance;
    return f;
}

// --------------------------------------------------------------------- OBJ --

namespace {

// Number parsing with the arithmetic of tinyobjloader's tryParseDouble
// (Part_2/3rdparty/tinyobjloader/tiny_obj_loader.h:897-1030): the mantissa is
// summed digit by digit in double (fraction digit k weighted by 10^-k from a
// table for k < 8), and an exponent e is applied as ldexp(m * 5^e, e).  The
// double is then rounded to float.  Keeping this arithmetic makes vertex
// positions bit-identical to the reference's mesh loader.
bool obj_parse_double(const char* s, const char* e, double* out) {
    static const double lut[] = {1.0, 0.1, 0.01, 0.001, 0.0001, 0.00001, 0.000001, 0.0000001};
    if (s >= e) return false;
    double mant = 0.0;
    int expo = 0;
    char sign = '+', esign = '+';
    const char* c = s;
    bool lead_dot = false;
    int read = 0;
    if (*c == '+' || *c == '-') {
        sign = *c++;
        if (c != e && *c == '.') lead_dot = true;
    } else if (*c >= '0' && *c <= '9') {
    } else if (*c == '.') {
        lead_dot = true;
    } else {
        return false;
    }
    if (!lead_dot) {
        while (c != e && *c >= '0' && *c <= '9') { mant = mant * 10 + (*c - '0'); ++c; ++read; }
        if (read == 0) return false;
    }
    if (c != e) {
        if (*c == '.') {
            ++c;
            read = 1;
            while (c != e && *c >= '0' && *c <= '9') {
                mant += (*c - '0') * (read < 8 ? lut[read] : std::pow(10.0, -read));
                ++read;
                ++c;
            }
        }
        if (c != e && (*c == 'e' || *c == 'E')) {
            ++c;
            if (c != e && (*c == '+' || *c == '-')) esign = *c++;
            else if (c == e || !(*c >= '0' && *c <= '9')) return false;
            read = 0;
            while (c != e && *c >= '0' && *c <= '9') {
                if (expo > 2147483647 / 10) return false;
                expo = expo * 10 + (*c - '0');
                ++c;
                ++read;
            }
            if (esign == '-') expo = -expo;
            if (read == 0) return false;
        }
    }
    *out = (sign == '+' ? 1 : -1) * (expo ? std::ldexp(mant * std::pow(5.0, expo), expo) : mant);
    return true;
}

float obj_real(const char*& p, double dflt) {
    while (*p == ' ' || *p == '\t') ++p;
    const char* e = p;
    while (*e && *e != ' ' && *e != '\t' && *e != '\r' && *e != '\n') ++e;
    double v = dflt;
    if (e > p && obj_parse_double(p, e, &v)) {
    } else {
        v = dflt;
    }
    p = e;
    return static_cast<float>(v);
}

struct ObjIdx {
    int v, t, n;
};

struct ObjKey {
    int v, n, t;
    bool operator==(const ObjKey& o) const { return v == o.v && n == o.n && t == o.t; }
};
struct ObjKeyHash {
    size_t operator()(const ObjKey& k) const {
        size_t h = std::hash<int>()(k.v);
        h ^= std::hash<int>()(k.n) + 0x9e3779b9 + (h << 6) + (h >> 2);
        h ^= std::hash<int>()(k.t) + 0x9e3779b9 + (h << 6) + (h >> 2);
        return h;
    }
};

int fix_index(int idx, int n) {  // 1-based, negative relative
    if (idx > 0) return idx - 1;
    if (idx < 0) return n + idx;
    return -1;
}

}  // namespace

// io/mesh_loader.cpp:23-112 (with tinyobj's triangulation: quads split along
// the shorter diagonal, tiny_obj_loader.h:1508-1600; larger polygons fanned).
bool load_obj(const std::string& path, int material_id, const float xform[16], std::vector<pt_vertex>& out_v,
              std::vector<pt_face>& out_f, std::string* err) {
    std::ifstream in(path);
    if (!in.is_open()) {
        if (err) *err = "cannot open OBJ '" + path + "'";
        return false;
    }
    std::vector<float> pos, nrm, tex;
    std::vector<std::vector<ObjIdx>> polys;
    std::string line;
    while (std::getline(in, line)) {
        const char* p = line.c_str();
        while (*p == ' ' || *p == '\t') ++p;
        if (p[0] == 'v' && (p[1] == ' ' || p[1] == '\t')) {
            p += 2;
            float x = obj_real(p, 0.0), y = obj_real(p, 0.0), z = obj_real(p, 0.0);
            pos.push_back(x); pos.push_back(y); pos.push_back(z);
        } else if (p[0] == 'v' && p[1] == 'n' && (p[2] == ' ' || p[2] == '\t')) {
            p += 3;
            float x = obj_real(p, 0.0), y = obj_real(p, 0.0), z = obj_real(p, 0.0);
            nrm.push_back(x); nrm.push_back(y); nrm.push_back(z);
        } else if (p[0] == 'v' && p[1] == 't' && (p[2] == ' ' || p[2] == '\t')) {
            p += 3;
            float u = obj_real(p, 0.0), v = obj_real(p, 0.0);
            tex.push_back(u); tex.push_back(v);
        } else if (p[0] == 'f' && (p[1] == ' ' || p[1] == '\t')) {
            p += 2;
            std::vector<ObjIdx> poly;
            int nv = static_cast<int>(pos.size() / 3), nt = static_cast<int>(tex.size() / 2),
                nn = static_cast<int>(nrm.size() / 3);
            while (*p) {
                while (*p == ' ' || *p == '\t') ++p;
                if (!*p || *p == '\r' || *p == '\n') break;
                ObjIdx ix{-1, -1, -1};
                ix.v = fix_index(std::atoi(p), nv);
                while (*p && *p != '/' && *p != ' ' && *p != '\t' && *p != '\r') ++p;
                if (*p == '/') {
                    ++p;
                    if (*p != '/') {
                        ix.t = fix_index(std::atoi(p), nt);
                        while (*p && *p != '/' && *p != ' ' && *p != '\t' && *p != '\r') ++p;
                    }
                    if (*p == '/') {
                        ++p;
                        ix.n = fix_index(std::atoi(p), nn);
                        while (*p && *p != ' ' && *p != '\t' && *p != '\r') ++p;
                    }
                }
                poly.push_back(ix);
            }
            if (poly.size() >= 3) polys.push_back(poly);
        }
    }

    float inv[16], invT[16];
    mat4_inverse(xform, inv);
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) invT[i * 4 + j] = inv[j * 4 + i];

    std::unordered_map<ObjKey, uint32_t, ObjKeyHash> vmap;
    auto get_vertex = [&](const ObjIdx& ix) -> uint32_t {
        ObjKey key{ix.v, ix.n, ix.t};
        auto it = vmap.find(key);
        if (it != vmap.end()) return it->second;
        pt_vertex vt;
        std::memset(&vt, 0, sizeof(vt));
        V3 P{pos[3 * ix.v + 0], pos[3 * ix.v + 1], pos[3 * ix.v + 2]};
        vt.position = to3p(xform_point(xform, P));
        if (ix.n >= 0) {
            V3 N{nrm[3 * ix.n + 0], nrm[3 * ix.n + 1], nrm[3 * ix.n + 2]};
            vt.normal = to3p(normalize(xform_vector(invT, N)));
        }
        if (ix.t >= 0) {
            vt.u = tex[2 * ix.t + 0];
            vt.v = tex[2 * ix.t + 1];
        }
        uint32_t id = static_cast<uint32_t>(out_v.size());
        out_v.push_back(vt);
        vmap[key] = id;
        return id;
    };
    auto sqdist = [&](int a, int b) {
        float dx = pos[3 * b + 0] - pos[3 * a + 0], dy = pos[3 * b + 1] - pos[3 * a + 1], dz = pos[3 * b + 2] - pos[3 * a + 2];
        return dx * dx + dy * dy + dz * dz;
    };
    auto emit = [&](const ObjIdx& a, const ObjIdx& b, const ObjIdx& c) {
        uint32_t i0 = get_vertex(a), i1 = get_vertex(b), i2 = get_vertex(c);
        if (a.n < 0) {  // mesh_loader.cpp:89-97: flat normal written to all three vertices
            V3 p0 = from(out_v[i0].position), p1 = from(out_v[i1].position), p2 = from(out_v[i2].position);
            pt_vec3p fn = to3p(normalize(cross(sub(p1, p0), sub(p2, p0))));
            out_v[i0].normal = fn; out_v[i1].normal = fn; out_v[i2].normal = fn;
        }
        out_f.push_back({i0, i1, i2, material_id});
    };
    int nv = static_cast<int>(pos.size() / 3);
    for (const auto& poly : polys) {
        bool ok = true;
        for (const auto& ix : poly) ok = ok && ix.v >= 0 && ix.v < nv;
        if (!ok) continue;
        if (poly.size() == 3) {
            emit(poly[0], poly[1], poly[2]);
        } else if (poly.size() == 4) {
            if (sqdist(poly[0].v, poly[2].v) < sqdist(poly[1].v, poly[3].v)) {
                emit(poly[0], poly[1], poly[2]);
                emit(poly[0], poly[2], poly[3]);
            } else {
                emit(poly[0], poly[1], poly[3]);
                emit(poly[1], poly[2], poly[3]);
            }
        } else {
            for (size_t k = 1; k + 1 < poly.size(); ++k) emit(poly[0], poly[k], poly[k + 1]);
        }
    }
    return true;
}

// ------------------------------------------------------------------ RGBE ---

// Radiance RGBE decode with the conversion of stb_image 2.30
// (Part_2/3rdparty/stb/stb_image.h:7131-7153): value = mantissa * 2^(e-136),
// e == 0 -> black; alpha = 1.  Flat and new-style run-length scanlines.
bool load_hdr(const std::string& path, Image& img, std::string* err) {
    std::ifstream in(path, std::ios::binary);
    if (!in.is_open()) {
        if (err) *err = "cannot open HDR '" + path + "'";
        return false;
    }
    std::string data((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    size_t p = 0;
    auto getline_ = [&](std::string& l) -> bool {
        if (p >= data.size()) return false;
        size_t e = data.find('\n', p);
        if (e == std::string::npos) e = data.size();
        l = data.substr(p, e - p);
        p = e + 1;
        return true;
    };
    std::string line;
    if (!getline_(line) || (line != "#?RADIANCE" && line != "#?RGBE")) {
        if (err) *err = "not a Radiance HDR file: '" + path + "'";
        return false;
    }
    bool fmt_ok = false;
    while (getline_(line)) {
        if (line.empty()) break;
        if (line == "FORMAT=32-bit_rle_rgbe") fmt_ok = true;
    }
    if (!fmt_ok) {
        if (err) *err = "unsupported HDR format in '" + path + "'";
        return false;
    }
    if (!getline_(line)) {
        if (err) *err = "missing HDR resolution line";
        return false;
    }
    int H = 0, W = 0;
    if (std::sscanf(line.c_str(), "-Y %d +X %d", &H, &W) != 2 || W <= 0 || H <= 0 || W > (1 << 24) ||
        H > (1 << 24)) {
        if (err) *err = "unsupported HDR orientation/size: '" + line + "'";
        return false;
    }
    img.width = W;
    img.height = H;
    img.rgba.assign(static_cast<size_t>(W) * H * 4, 0.0f);
    auto put = [&](size_t idx, const unsigned char* rgbe) {
        float* o = &img.rgba[idx * 4];
        if (rgbe[3] != 0) {
            float f = static_cast<float>(std::ldexp(1.0f, rgbe[3] - (128 + 8)));
            o[0] = rgbe[0] * f; o[1] = rgbe[1] * f; o[2] = rgbe[2] * f;
        } else {
            o[0] = o[1] = o[2] = 0.0f;
        }
        o[3] = 1.0f;
    };
    auto avail = [&](size_t n) { return p + n <= data.size(); };
    const unsigned char* u = reinterpret_cast<const unsigned char*>(data.data());
    auto flat_from = [&](size_t first) -> bool {
        size_t total = static_cast<size_t>(W) * H;
        for (size_t i = first; i < total; ++i) {
            if (!avail(4)) return false;
            put(i, u + p);
            p += 4;
        }
        return true;
    };
    bool ok = true;
    if (W < 8 || W >= 32768) {
        ok = flat_from(0);
    } else {
        std::vector<unsigned char> scan(static_cast<size_t>(W) * 4);
        for (int y = 0; y < H && ok; ++y) {
            if (!avail(4)) { ok = false; break; }
            unsigned char c1 = u[p], c2 = u[p + 1], hi = u[p + 2], lo = u[p + 3];
            if (c1 != 2 || c2 != 2 || (hi & 0x80)) {
                // not run-length encoded: decode the rest of the file flat
                ok = flat_from(static_cast<size_t>(y) * W);
                break;
            }
            p += 4;
            int len = (hi << 8) | lo;
            if (len != W) { ok = false; break; }
            for (int ch = 0; ch < 4 && ok; ++ch) {
                int x = 0;
                while (x < W) {
                    if (!avail(1)) { ok = false; break; }
                    int count = u[p++];
                    if (count > 128) {
                        count -= 128;
                        if (count > W - x || !avail(1)) { ok = false; break; }
                        unsigned char v = u[p++];
                        for (int k = 0; k < count; ++k) scan[(x++) * 4 + ch] = v;
                    } else {
                        if (count == 0 || count > W - x || !avail(count)) { ok = false; break; }
                        for (int k = 0; k < count; ++k) scan[(x++) * 4 + ch] = u[p++];
                    }
                }
            }
            for (int x = 0; x < W && ok; ++x) put(static_cast<size_t>(y) * W + x, &scan[x * 4]);
        }
    }
    if (!ok) {
        if (err) *err = "corrupt HDR data in '" + path + "'";
        img = Image();
        return false;
    }
    return true;
}

// Radiance writer (flat scanlines).  flip_y: film row 0 is the bottom row, as
// ImageIO::write flips (io/image_io.cpp:110-117).
bool write_hdr(const std::string& path, int w, int h, const float* rgba, bool flip_y, std::string* err) {
    std::ofstream out(path, std::ios::binary);
    if (!out.is_open()) {
        if (err) *err = "cannot write '" + path + "'";
        return false;
    }
    out << "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y " << h << " +X " << w << "\n";
    std::vector<unsigned char> row(static_cast<size_t>(w) * 4);
    for (int yy = 0; yy < h; ++yy) {
        int y = flip_y ? (h - 1 - yy) : yy;
        for (int x = 0; x < w; ++x) {
            const float* c = rgba + (static_cast<size_t>(y) * w + x) * 4;
            float r = std::fmax(c[0], 0.0f), g = std::fmax(c[1], 0.0f), b = std::fmax(c[2], 0.0f);
            float m = std::fmax(r, std::fmax(g, b));
            unsigned char* o = &row[x * 4];
            if (!(m >= 1e-32f)) {
                o[0] = o[1] = o[2] = o[3] = 0;
            } else {
                int e;
                float mant = std::frexp(m, &e);
                float s = mant * 256.0f / m;
                o[0] = static_cast<unsigned char>(r * s);
                o[1] = static_cast<unsigned char>(g * s);
                o[2] = static_cast<unsigned char>(b * s);
                o[3] = static_cast<unsigned char>(e + 128);
            }
        }
        out.write(reinterpret_cast<const char*>(row.data()), static_cast<std::streamsize>(row.size()));
    }
    return static_cast<bool>(out);
}

bool write_pfm(const std::string& path, int w, int h, const float* rgba, std::string* err) {
    std::ofstream out(path, std::ios::binary);
    if (!out.is_open()) {
        if (err) *err = "cannot write '" + path + "'";
        return false;
    }
    out << "PF\n" << w << " " << h << "\n-1.0\n";
    std::vector<float> row(static_cast<size_t>(w) * 3);
    for (int y = 0; y < h; ++y) {  // PFM stores the bottom row first, like the film
        for (int x = 0; x < w; ++x)
            for (int c = 0; c < 3; ++c) row[x * 3 + c] = rgba[(static_cast<size_t>(y) * w + x) * 4 + c];
        out.write(reinterpret_cast<const char*>(row.data()), static_cast<std::streamsize>(row.size() * 4));
    }
    return static_cast<bool>(out);
}

// ----------------------------------------------------------------- ImageIO --

namespace {
std::string lower_ext(const std::string& path) {
    const size_t dot = path.find_last_of('.');
    std::string ext = dot == std::string::npos ? std::string() : path.substr(dot);
    for (auto& ch : ext) ch = static_cast<char>(std::tolower(static_cast<unsigned char>(ch)));
    return ext;
}
}  // namespace

// ImageIO::load (io/image_io.cpp:17-67): .hdr as float RGB, .png/.jpg/.jpeg
// as 8-bit RGB decoded with pow(v / 255, 2.2); alpha is 1 in both.
bool load_image(const std::string& path, Image& img, std::string* err) {
    const std::string ext = lower_ext(path);
    if (ext == ".hdr") return load_hdr(path, img, err);
    if (ext != ".png" && ext != ".jpg" && ext != ".jpeg") {
        if (err) *err = "unsupported image format: '" + path + "'";
        return false;
    }
    std::vector<uint8_t> bytes;
    Image8 im;
    std::string derr;
    if (!read_file(path, bytes, err)) return false;
    if (!decode_image(bytes.data(), bytes.size(), im, &derr)) {
        if (err) *err = "'" + path + "': " + derr;
        return false;
    }
    img.width = im.width;
    img.height = im.height;
    img.rgba.resize(static_cast<size_t>(im.width) * im.height * 4);
    for (size_t i = 0; i < static_cast<size_t>(im.width) * im.height; ++i) {
        for (int c = 0; c < 3; ++c) img.rgba[i * 4 + c] = std::pow(im.rgba[i * 4 + c] / 255.0f, 2.2f);
        img.rgba[i * 4 + 3] = 1.0f;
    }
    return true;
}

// The 8-bit value ImageIO::write stores for a linear channel
// (io/image_io.cpp:119-127): clamp(pow(v, 1/2.2), 0, 1) * 255 + 0.5, truncated.
// A NaN (the power of a negative value) is stored as 0.
uint8_t ldr_byte(float v) {
    float g = std::pow(v, 1.0f / 2.2f);
    if (!(g >= 0.0f)) g = 0.0f;
    if (g > 1.0f) g = 1.0f;
    return static_cast<uint8_t>(g * 255.0f + 0.5f);
}

// ImageIO::write (io/image_io.cpp:79-144) for .hdr, .pfm, .png, .jpg/.jpeg
// (quality 90), .bmp and .tga; flip_y writes row h-1 of rgba first.
bool write_image(const std::string& path, int w, int h, const float* rgba, bool flip_y, std::string* err) {
    const std::string ext = lower_ext(path);
    if (ext == ".hdr") return write_hdr(path, w, h, rgba, flip_y, err);
    if (ext == ".pfm") return write_pfm(path, w, h, rgba, err);
    if (ext != ".png" && ext != ".jpg" && ext != ".jpeg" && ext != ".bmp" && ext != ".tga") {
        if (err) *err = "unsupported image format: '" + path + "'";
        return false;
    }
    std::vector<uint8_t> rgb(static_cast<size_t>(w) * h * 3), out;
    for (int y = 0; y < h; ++y) {
        const float* src = rgba + static_cast<size_t>(flip_y ? h - 1 - y : y) * w * 4;
        for (int x = 0; x < w; ++x)
            for (int c = 0; c < 3; ++c) rgb[(static_cast<size_t>(y) * w + x) * 3 + c] = ldr_byte(src[x * 4 + c]);
    }
    if (ext == ".png") encode_png(w, h, rgb.data(), out);
    else if (ext == ".bmp") encode_bmp(w, h, rgb.data(), out);
    else if (ext == ".tga") encode_tga(w, h, rgb.data(), out);
    else encode_jpeg(w, h, rgb.data(), 90, out);
    std::ofstream f(path, std::ios::binary);
    if (!f.is_open()) {
        if (err) *err = "cannot write '" + path + "'";
        return false;
    }
    f.write(reinterpret_cast<const char*>(out.data()), static_cast<std::streamsize>(out.size()));
    if (!f) {
        if (err) *err = "cannot write '" + path + "'";
        return false;
    }
    return true;
}

// -------------------------------------------------------------- JSON scene --

namespace {

std::string resolve_relative(const std::string& base, const std::string& rel) {
    if (!rel.empty() && rel[0] == '/') return rel;
    size_t s = base.find_last_of('/');
    std::string dir = (s == std::string::npos) ? std::string() : base.substr(0, s + 1);
    return dir + rel;
}

void to_vec3(const Json& j, float out[3]) {
    out[0] = j[0].as_float();
    out[1] = j[1].as_float();
    out[2] = j[2].as_float();
}

// scene_loader.cpp:34-52
void apply_material_override(const Json& p, pt_material& m) {
    float v[3];
    if (p.contains("baseColor")) { to_vec3(p["baseColor"], v); m.base_color = {v[0], v[1], v[2], 0.0f}; }
    if (p.contains("metallic")) m.metallic = p["metallic"].as_float();
    if (p.contains("roughness")) m.roughness = p["roughness"].as_float();
    if (p.contains("specularTint")) m.specular_tint = p["specularTint"].as_float();
    if (p.contains("anisotropic")) m.anisotropic = p["anisotropic"].as_float();
    if (p.contains("sheen")) m.sheen = p["sheen"].as_float();
    if (p.contains("sheenTint")) m.sheen_tint = p["sheenTint"].as_float();
    if (p.contains("clearcoat")) m.clearcoat = p["clearcoat"].as_float();
    if (p.contains("clearcoatRoughness")) m.clearcoat_roughness = p["clearcoatRoughness"].as_float();
    if (p.contains("specTrans")) m.spec_trans = p["specTrans"].as_float();
    if (p.contains("ior")) m.ior = p["ior"].as_float();
    if (p.contains("subsurface")) m.subsurface = p["subsurface"].as_float();
    if (p.contains("emissive")) { to_vec3(p["emissive"], v); m.emissive = {v[0], v[1], v[2], 0.0f}; }
    if (p.contains("emissiveStrength")) m.emissive_strength = p["emissiveStrength"].as_float();
    if (p.contains("alphaCutoff")) m.alpha_cutoff = p["alphaCutoff"].as_float();
    if (p.contains("alphaMode")) m.alpha_mode = p["alphaMode"].as_int();
    if (p.contains("doubleSided")) m.double_sided = p["doubleSided"].as_bool() ? 1 : 0;
}

// scene_loader.cpp:55-73
void parse_transform(const Json& obj, float out[16]) {
    mat4_identity(out);
    if (!obj.contains("transform")) return;
    const Json& t = obj["transform"];
    float tr[3] = {0, 0, 0}, axis[3] = {0, 1, 0};
    float angle = 0.0f, scale = 1.0f;
    if (t.contains("translate")) to_vec3(t["translate"], tr);
    if (t.contains("rotate")) {
        to_vec3(t["rotate"]["axis"], axis);
        angle = t["rotate"]["angle"].as_float();
    }
    if (t.contains("scale")) scale = t["scale"].as_float();
    mat4_trs(tr, axis, angle, scale, out);
}

}  // namespace

// SceneLoader::load (scene_loader.cpp:278-292): by extension.
bool load_scene(const std::string& path, HostScene& scene, std::string* err) {
    const std::string ext = lower_ext(path);
    if (ext == ".json") return load_scene_json(path, scene, err);
    if (ext == ".gltf" || ext == ".glb") {
        init_scene_defaults(scene);
        scene.path = path;
        return load_gltf(path, scene, GltfOptions{}, err);
    }
    if (err) *err = "unknown scene format: " + path;
    return false;
}

// SceneLoader::loadFromJson, scene_loader.cpp:240-276.  Extensions (ignored
// by the reference, which skips unknown keys): render.seed, environment
// "generate", material "albedoTexture" (a linear float RGB image), objects of
// type "sphere" (icosphere), and a top-level "bvh" block.
bool load_scene_json(const std::string& path, HostScene& scene, std::string* err) {
    std::ifstream f(path);
    if (!f.is_open()) {
        if (err) *err = "cannot open " + path;
        return false;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    Json root;
    std::string perr;
    if (!json_parse(ss.str(), root, &perr)) {
        if (err) *err = "JSON parse error in " + path + ": " + perr;
        return false;
    }
    if (!root.is_object()) {
        if (err) *err = "scene root is not an object: " + path;
        return false;
    }
    scene.path = path;
    try {
        bool camera_set = false;
        if (root.contains("camera")) {  // :75-89
            const Json& c = root["camera"];
            float v[3];
            if (c.contains("position")) { to_vec3(c["position"], v); scene.camera.position = {v[0], v[1], v[2], 0}; }
            if (c.contains("lookAt")) { to_vec3(c["lookAt"], v); scene.camera.look_at = {v[0], v[1], v[2], 0}; }
            if (c.contains("up")) { to_vec3(c["up"], v); scene.camera.up = {v[0], v[1], v[2], 0}; }
            if (c.contains("fov")) scene.camera.fov = c["fov"].as_float();
            if (c.contains("aperture")) scene.camera.aperture = c["aperture"].as_float();
            if (c.contains("focalDistance")) scene.camera.focal_distance = c["focalDistance"].as_float();
            if (c.contains("resolution")) {
                scene.camera.width = c["resolution"][0].as_int();
                scene.camera.height = c["resolution"][1].as_int();
            }
            camera_set = true;
        }
        if (root.contains("render")) {  // :91-103
            const Json& r = root["render"];
            if (r.contains("spp")) scene.settings.target_samples = r["spp"].as_int();
            if (r.contains("samplePerFrame")) scene.settings.sample_per_frame = r["samplePerFrame"].as_int();
            if (r.contains("maxDepth")) scene.settings.max_depth = r["maxDepth"].as_int();
            if (r.contains("output")) scene.output_file = r["output"].str;
            if (r.contains("sampler")) {
                const std::string& s = r["sampler"].str;
                if (s == "pcg") scene.settings.sampler = 0;
                else if (s == "sobol") scene.settings.sampler = 1;
            }
            if (r.contains("seed")) scene.settings.seed = static_cast<uint32_t>(r["seed"].num);
        }
        if (root.contains("environment")) {  // :105-110
            const Json& e = root["environment"];
            if (e.contains("file")) {
                std::string hp = resolve_relative(path, e["file"].str);
                std::string herr;
                if (!load_image(hp, scene.env, &herr)) {
                    // ImageIO::load failure leaves an empty image; the reference renders black
                    // (device_scene.cu:19-27).  Report it instead of rendering silently.
                    if (err) *err = herr;
                    return false;
                }
            } else if (e.contains("generate")) {
                const Json& g = e["generate"];
                int w = g.contains("width") ? g["width"].as_int() : 2048;
                int h = g.contains("height") ? g["height"].as_int() : 1024;
                uint64_t seed = g.contains("seed") ? static_cast<uint64_t>(g["seed"].num) : 0;
                generate_sky_env(scene.env, w, h, seed);
            }
        }
        std::map<std::string, int> mat_index;
        if (root.contains("materials")) {  // :112-122 (keys in sorted order)
            for (const auto& kv : root["materials"].obj) {
                pt_material m = default_material();
                apply_material_override(kv.second, m);
                if (kv.second.contains("albedoTexture")) {
                    const Json& t = kv.second["albedoTexture"];
                    Image img;
                    img.width = t["width"].as_int();
                    img.height = t["height"].as_int();
                    img.linear = 1;
                    float fill[3] = {1.0f, 1.0f, 1.0f};
                    if (t.contains("fill")) to_vec3(t["fill"], fill);
                    if (img.width <= 0 || img.height <= 0 || img.width > 16384 || img.height > 16384) {
                        if (err) *err = "albedoTexture: bad size";
                        return false;
                    }
                    img.rgba.resize(static_cast<size_t>(img.width) * img.height * 4);
                    for (size_t i = 0; i < static_cast<size_t>(img.width) * img.height; ++i) {
                        img.rgba[i * 4 + 0] = fill[0]; img.rgba[i * 4 + 1] = fill[1];
                        img.rgba[i * 4 + 2] = fill[2]; img.rgba[i * 4 + 3] = 1.0f;
                    }
                    m.base_color_tex_id = static_cast<int>(scene.textures.size());
                    scene.textures.push_back(std::move(img));
                }
                mat_index[kv.first] = static_cast<int>(scene.materials.size());
                scene.materials.push_back(m);
            }
        }
        if (root.contains("lights")) {  // :124-142
            for (const auto& jl : root["lights"].arr) {
                pt_punctual_light l = default_light();
                std::string type = jl.contains("type") ? jl["type"].str : std::string("point");
                if (type == "directional") l.type = 1;
                else if (type == "spot") l.type = 2;
                else l.type = 0;
                float v[3];
                if (jl.contains("position")) { to_vec3(jl["position"], v); l.position = {v[0], v[1], v[2], 0}; }
                if (jl.contains("direction")) {
                    to_vec3(jl["direction"], v);
                    V3 d = normalize(V3{v[0], v[1], v[2]});
                    l.direction = {d.x, d.y, d.z, 0};
                }
                if (jl.contains("color")) { to_vec3(jl["color"], v); l.color = {v[0], v[1], v[2], 0}; }
                if (jl.contains("intensity")) l.intensity = jl["intensity"].as_float();
                if (jl.contains("range")) l.range = jl["range"].as_float();
                if (jl.contains("innerCone")) l.inner_cone = std::cos(jl["innerCone"].as_float());
                if (jl.contains("outerCone")) l.outer_cone = std::cos(jl["outerCone"].as_float());
                scene.lights.push_back(l);
            }
        }
        if (root.contains("objects")) {  // :144-209
            auto resolve_mat = [&](const Json& j) -> int {
                if (j.is_string()) {
                    auto it = mat_index.find(j.str);
                    return it != mat_index.end() ? it->second : 0;
                }
                return j.as_int();
            };
            for (const auto& obj : root["objects"].arr) {
                std::string type = obj.contains("type") ? obj["type"].str : std::string();
                if (type == "quad") {
                    const Json& vs = obj["vertices"];
                    float a[3], b[3], c[3], d[3];
                    to_vec3(vs[0], a); to_vec3(vs[1], b); to_vec3(vs[2], c); to_vec3(vs[3], d);
                    build_quad(a, b, c, d, resolve_mat(obj["material"]), scene.vertices, scene.faces);
                } else if (type == "mesh") {
                    std::string mp = resolve_relative(path, obj["file"].str);
                    int mat = resolve_mat(obj["material"]);
                    float xf[16];
                    parse_transform(obj, xf);
                    std::string merr;
                    if (!load_obj(mp, mat, xf, scene.vertices, scene.faces, &merr)) {
                        // the reference prints the loader error and keeps going (mesh_loader.cpp:35-38)
                        if (err) *err = merr;
                        return false;
                    }
                } else if (type == "gltf") {  // :162-194
                    const std::string gp = resolve_relative(path, obj["file"].str);
                    GltfOptions go;
                    parse_transform(obj, go.transform);
                    go.load_camera = go.auto_fit_camera = go.add_default_lights = go.load_environment = false;
                    std::vector<std::string> names;
                    go.material_names = &names;
                    const size_t first = scene.materials.size();
                    std::string gerr;
                    if (!load_gltf(gp, scene, go, &gerr)) {
                        // the reference prints the loader error and keeps going (gltf_loader.cpp:600-604)
                        if (err) *err = gerr;
                        return false;
                    }
                    if (obj.contains("materialOverrides"))
                        for (const auto& kv : obj["materialOverrides"].obj)
                            for (size_t i = 0; i < names.size(); ++i)
                                if (names[i] == kv.first && first + i < scene.materials.size())
                                    apply_material_override(kv.second, scene.materials[first + i]);
                }
            }
        }
        if (root.contains("bvh")) {
            const Json& b = root["bvh"];
            scene.bvh_opts.bins = b.contains("bins") ? b["bins"].as_int() : 24;
            scene.bvh_opts.leaf_threshold = b.contains("leafThreshold") ? b["leafThreshold"].as_int() : 4;
            scene.bvh_opts.max_leaf = b.contains("maxLeaf") ? b["maxLeaf"].as_int() : 0;
            scene.has_bvh_opts = true;
        }
        if (scene.materials.empty()) scene.materials.push_back(default_material());  // :265-266
        if (!camera_set) auto_fit_camera(scene.vertices, scene.camera);             // :268-269
    } catch (const std::exception& e) {
        if (err) *err = std::string("scene '") + path + "': " + e.what();
        return false;
    }
    return true;
}

}  // namespace ptamd

// glTF 2.0 scenes (.gltf with external or data-URI buffers, .glb) into a
// HostScene, with the semantics of the reference's GltfSceneLoader
// (Part_2/src/io/gltf_loader.cpp): the default scene's node tree with TRS or
// matrix transforms, triangle primitives (float positions / normals / first
// uv set / tangents, u8/u16/u32 indices), PBR metallic-roughness materials
// with the KHR material extensions the reference reads, KHR_texture_transform
// on the base colour texture, textures decoded to RGBA8 with their sampler's
// wrap modes, the first perspective camera, KHR_lights_punctual, and the
// fallbacks (camera fitted to the bounds, default environment, default
// lights).  JSON comes from json.cpp, images from image_codec.cpp.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <filesystem>
#include <stdexcept>

#include "gltf.h"
#include "image_codec.h"
#include "json.h"

namespace ptamd {

namespace {

struct V3 {
    float x, y, z;
};
V3 xf_point(const float m[16], V3 v) {
    return {m[0] * v.x + m[1] * v.y + m[2] * v.z + m[3], m[4] * v.x + m[5] * v.y + m[6] * v.z + m[7],
            m[8] * v.x + m[9] * v.y + m[10] * v.z + m[11]};
}
V3 xf_vector(const float m[16], V3 v) {
    return {m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
            m[8] * v.x + m[9] * v.y + m[10] * v.z};
}
V3 normalized(V3 v) {
    const float l = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    return l > 0.0f ? V3{v.x / l, v.y / l, v.z / l} : v;
}
pt_vec3p p3(V3 v) { return pt_vec3p{v.x, v.y, v.z, 0.0f}; }

double num(const Json& j, const char* k, double def) {
    if (!j.contains(k)) return def;
    const Json& v = j[k];
    if (v.type == Json::Number) return v.num;
    if (v.type == Json::Bool) return v.b ? 1.0 : 0.0;
    return def;
}
int inum(const Json& j, const char* k, int def) { return j.contains(k) && j[k].is_number() ? static_cast<int>(j[k].num) : def; }
const Json* member(const Json& j, const char* k) { return j.contains(k) ? &j[k] : nullptr; }
bool arr3(const Json* a, float out[3]) {
    if (!a || !a->is_array() || a->size() < 3) return false;
    for (int i = 0; i < 3; ++i) out[i] = (*a)[static_cast<size_t>(i)].as_float();
    return true;
}

std::string uri_decode(const std::string& s) {
    std::string o;
    for (size_t i = 0; i < s.size(); ++i) {
        if (s[i] == '%' && i + 2 < s.size() && std::isxdigit(static_cast<unsigned char>(s[i + 1])) &&
            std::isxdigit(static_cast<unsigned char>(s[i + 2]))) {
            o.push_back(static_cast<char>(std::stoi(s.substr(i + 1, 2), nullptr, 16)));
            i += 2;
        } else {
            o.push_back(s[i]);
        }
    }
    return o;
}

bool base64_decode(const std::string& in, std::vector<uint8_t>& out) {
    auto val = [](char c) -> int {
        if (c >= 'A' && c <= 'Z') return c - 'A';
        if (c >= 'a' && c <= 'z') return c - 'a' + 26;
        if (c >= '0' && c <= '9') return c - '0' + 52;
        if (c == '+' || c == '-') return 62;
        if (c == '/' || c == '_') return 63;
        return -1;
    };
    out.clear();
    uint32_t acc = 0;
    int bits = 0;
    for (char c : in) {
        if (c == '=' || c == '\n' || c == '\r' || c == ' ') continue;
        const int v = val(c);
        if (v < 0) return false;
        acc = (acc << 6) | static_cast<uint32_t>(v);
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            out.push_back(static_cast<uint8_t>((acc >> bits) & 0xFFu));
        }
    }
    return true;
}

// A buffer or image URI: a data URI or a file relative to the glTF file.
bool load_uri(const std::string& uri, const std::string& base_dir, std::vector<uint8_t>& out, std::string* err) {
    if (uri.compare(0, 5, "data:") == 0) {
        const size_t comma = uri.find(',');
        if (comma == std::string::npos || uri.substr(0, comma).find(";base64") == std::string::npos) {
            if (err) *err = "glTF: only base64 data URIs are supported";
            return false;
        }
        if (!base64_decode(uri.substr(comma + 1), out)) {
            if (err) *err = "glTF: bad base64 data";
            return false;
        }
        return true;
    }
    return read_file(base_dir + uri_decode(uri), out, err);
}

// The glTF document with its buffers.
struct Doc {
    Json j;
    std::vector<std::vector<uint8_t>> buffers;
    std::string dir;
    const Json& list(const char* k) const {
        static const Json empty;
        return j.contains(k) && j[k].is_array() ? j[k] : empty;
    }
};

// An accessor's elements as floats (normalized integer types scaled to
// [0, 1] / [-1, 1] as the glTF spec defines), stride-aware.
struct Accessor {
    const uint8_t* base = nullptr;
    size_t stride = 0, count = 0;
    int ctype = 5126, ncomp = 1;
    bool normalized = false;
    float get(size_t i, int c) const {
        const uint8_t* p = base + i * stride;
        switch (ctype) {
            case 5126: {
                float f;
                std::memcpy(&f, p + 4 * c, 4);
                return f;
            }
            case 5121: return normalized ? p[c] / 255.0f : static_cast<float>(p[c]);
            case 5120: {
                const float v = static_cast<float>(static_cast<int8_t>(p[c]));
                return normalized ? std::max(v / 127.0f, -1.0f) : v;
            }
            case 5123: {
                uint16_t v;
                std::memcpy(&v, p + 2 * c, 2);
                return normalized ? v / 65535.0f : static_cast<float>(v);
            }
            case 5122: {
                int16_t v;
                std::memcpy(&v, p + 2 * c, 2);
                return normalized ? std::max(v / 32767.0f, -1.0f) : static_cast<float>(v);
            }
            case 5125: {
                uint32_t v;
                std::memcpy(&v, p + 4 * c, 4);
                return static_cast<float>(v);
            }
        }
        return 0.0f;
    }
    uint32_t index(size_t i) const {
        const uint8_t* p = base + i * stride;
        if (ctype == 5121) return p[0];
        if (ctype == 5123) {
            uint16_t v;
            std::memcpy(&v, p, 2);
            return v;
        }
        uint32_t v;
        std::memcpy(&v, p, 4);
        return v;
    }
};

int comp_bytes(int ctype) {
    switch (ctype) {
        case 5120: case 5121: return 1;
        case 5122: case 5123: return 2;
        case 5125: case 5126: return 4;
    }
    return 0;
}
int type_comps(const std::string& t) {
    if (t == "SCALAR") return 1;
    if (t == "VEC2") return 2;
    if (t == "VEC3") return 3;
    if (t == "VEC4") return 4;
    if (t == "MAT4") return 16;
    return 0;
}

Accessor accessor(const Doc& d, int idx) {
    const Json& accs = d.list("accessors");
    if (idx < 0 || static_cast<size_t>(idx) >= accs.size()) throw std::runtime_error("glTF: accessor index out of range");
    const Json& a = accs[static_cast<size_t>(idx)];
    Accessor r;
    r.count = static_cast<size_t>(num(a, "count", 0));
    r.ctype = inum(a, "componentType", 5126);
    r.ncomp = type_comps(a.contains("type") ? a["type"].str : std::string("SCALAR"));
    r.normalized = a.contains("normalized") && a["normalized"].as_bool();
    const int cb = comp_bytes(r.ctype);
    if (cb == 0 || r.ncomp == 0) throw std::runtime_error("glTF: unsupported accessor type");
    if (!a.contains("bufferView")) throw std::runtime_error("glTF: accessors without a bufferView are not supported");
    const Json& views = d.list("bufferViews");
    const int vi = a["bufferView"].as_int();
    if (vi < 0 || static_cast<size_t>(vi) >= views.size()) throw std::runtime_error("glTF: bufferView index out of range");
    const Json& v = views[static_cast<size_t>(vi)];
    const int bi = inum(v, "buffer", 0);
    if (bi < 0 || static_cast<size_t>(bi) >= d.buffers.size()) throw std::runtime_error("glTF: buffer index out of range");
    const std::vector<uint8_t>& buf = d.buffers[static_cast<size_t>(bi)];
    const size_t off = static_cast<size_t>(num(v, "byteOffset", 0)) + static_cast<size_t>(num(a, "byteOffset", 0));
    const size_t elem = static_cast<size_t>(cb) * r.ncomp;
    r.stride = v.contains("byteStride") && v["byteStride"].as_int() > 0 ? static_cast<size_t>(v["byteStride"].as_int()) : elem;
    if (r.count > 0 && off + (r.count - 1) * r.stride + elem > buf.size())
        throw std::runtime_error("glTF: accessor reaches past its buffer");
    r.base = buf.data() + off;
    return r;
}

// Textures (gltf_loader.cpp:112-148): image decoded to RGBA8 -> floats / 255,
// the sampler's wrap modes.  A texture whose image is missing or cannot be
// decoded becomes a 1x1 texel of zeros (what fetching an unset texture object
// returns on the reference's device).
void load_textures(const Doc& d, std::vector<Image>& out, std::vector<std::string>* warnings) {
    const Json& texs = d.list("textures");
    const Json& imgs = d.list("images");
    const Json& samplers = d.list("samplers");
    const Json& views = d.list("bufferViews");
    std::vector<Image8> decoded(imgs.size());
    std::vector<bool> ok(imgs.size(), false);
    for (size_t i = 0; i < imgs.size(); ++i) {
        const Json& im = imgs[i];
        std::vector<uint8_t> bytes;
        std::string err;
        bool have = false;
        if (im.contains("bufferView")) {
            const int vi = im["bufferView"].as_int();
            if (vi >= 0 && static_cast<size_t>(vi) < views.size()) {
                const Json& v = views[static_cast<size_t>(vi)];
                const int bi = inum(v, "buffer", 0);
                const size_t off = static_cast<size_t>(num(v, "byteOffset", 0)), len = static_cast<size_t>(num(v, "byteLength", 0));
                if (bi >= 0 && static_cast<size_t>(bi) < d.buffers.size() && off + len <= d.buffers[static_cast<size_t>(bi)].size()) {
                    const auto& b = d.buffers[static_cast<size_t>(bi)];
                    bytes.assign(b.begin() + static_cast<std::ptrdiff_t>(off), b.begin() + static_cast<std::ptrdiff_t>(off + len));
                    have = true;
                }
            }
        } else if (im.contains("uri")) {
            have = load_uri(im["uri"].str, d.dir, bytes, &err);
        }
        if (have) have = decode_image(bytes.data(), bytes.size(), decoded[i], &err);
        ok[i] = have;
        if (!have && warnings) warnings->push_back("image " + std::to_string(i) + ": " + err);
    }
    out.resize(texs.size());
    for (size_t t = 0; t < texs.size(); ++t) {
        const Json& tx = texs[t];
        Image& o = out[t];
        const int src = inum(tx, "source", -1);
        if (src >= 0 && static_cast<size_t>(src) < imgs.size() && ok[static_cast<size_t>(src)]) {
            const Image8& im = decoded[static_cast<size_t>(src)];
            o.width = im.width;
            o.height = im.height;
            o.rgba.resize(im.rgba.size());
            for (size_t k = 0; k < im.rgba.size(); ++k) o.rgba[k] = im.rgba[k] / 255.0f;
        } else {
            o.width = o.height = 1;
            o.rgba.assign(4, 0.0f);
        }
        const int si = inum(tx, "sampler", -1);
        if (si >= 0 && static_cast<size_t>(si) < samplers.size()) {
            o.wrap_s = inum(samplers[static_cast<size_t>(si)], "wrapS", 10497);
            o.wrap_t = inum(samplers[static_cast<size_t>(si)], "wrapT", 10497);
        }
    }
}

int tex_index(const Json* info, int offset) {
    if (!info || !info->contains("index")) return -1;
    return (*info)["index"].as_int() + offset;
}

// Materials (gltf_loader.cpp:172-335).
void load_materials(const Doc& d, std::vector<pt_material>& out, int toff, std::vector<std::string>* names) {
    const Json& mats = d.list("materials");
    out.assign(mats.size(), default_material());
    if (names) names->assign(mats.size(), std::string());
    for (size_t i = 0; i < mats.size(); ++i) {
        const Json& g = mats[i];
        pt_material& m = out[i];
        if (names && g.contains("name")) (*names)[i] = g["name"].str;
        static const Json none;
        const Json& pbr = g.contains("pbrMetallicRoughness") ? g["pbrMetallicRoughness"] : none;
        // absent factors take the glTF defaults (base colour [1,1,1,1], emissive [0,0,0]),
        // which the reference copies like given ones (gltf_loader.cpp:185-193, 216-221)
        const Json* bcf = member(pbr, "baseColorFactor");
        if (bcf && bcf->is_array() && bcf->size() >= 3) {
            m.base_color = pt_vec3p{(*bcf)[0].as_float(), (*bcf)[1].as_float(), (*bcf)[2].as_float(), 0.0f};
            if (bcf->size() >= 4) m.base_alpha = (*bcf)[3].as_float();
        } else if (!bcf) {
            m.base_color = pt_vec3p{1.0f, 1.0f, 1.0f, 0.0f};
            m.base_alpha = 1.0f;
        }
        m.metallic = static_cast<float>(num(pbr, "metallicFactor", 1.0));    // glTF default 1
        m.roughness = static_cast<float>(num(pbr, "roughnessFactor", 1.0));  // glTF default 1
        if (const Json* bt = member(pbr, "baseColorTexture")) {
            m.base_color_tex_id = tex_index(bt, toff);
            if (const Json* ext = member(*bt, "extensions"))
                if (const Json* tt = member(*ext, "KHR_texture_transform")) {
                    const Json* o = member(*tt, "offset");
                    if (o && o->is_array() && o->size() >= 2) {
                        m.uv_offset_x = (*o)[0].as_float();
                        m.uv_offset_y = (*o)[1].as_float();
                    }
                    const Json* s = member(*tt, "scale");
                    if (s && s->is_array() && s->size() >= 2) {
                        m.uv_scale_x = (*s)[0].as_float();
                        m.uv_scale_y = (*s)[1].as_float();
                    }
                    if (tt->contains("rotation")) m.uv_rotation = (*tt)["rotation"].as_float();
                }
        }
        if (const Json* t = member(pbr, "metallicRoughnessTexture")) m.metallic_roughness_tex_id = tex_index(t, toff);
        if (const Json* t = member(g, "normalTexture")) {
            m.normal_tex_id = tex_index(t, toff);
            m.normal_scale = static_cast<float>(num(*t, "scale", 1.0));
        }
        if (const Json* t = member(g, "emissiveTexture")) m.emissive_tex_id = tex_index(t, toff);
        if (const Json* t = member(g, "occlusionTexture")) {
            m.occlusion_tex_id = tex_index(t, toff);
            m.occlusion_strength = static_cast<float>(num(*t, "strength", 1.0));
        }
        float ef[3] = {0.0f, 0.0f, 0.0f};
        if (arr3(member(g, "emissiveFactor"), ef) || !g.contains("emissiveFactor")) {
            m.emissive = pt_vec3p{ef[0], ef[1], ef[2], 0.0f};
            if (ef[0] > 0 || ef[1] > 0 || ef[2] > 0) m.emissive_strength = 1.0f;
        }
        m.double_sided = g.contains("doubleSided") && g["doubleSided"].as_bool() ? 1 : 0;
        if (const Json* ext = member(g, "extensions")) {
            if (const Json* e = member(*ext, "KHR_materials_emissive_strength"))
                if (e->contains("emissiveStrength")) m.emissive_strength = (*e)["emissiveStrength"].as_float();
            if (const Json* e = member(*ext, "KHR_materials_ior"))
                if (e->contains("ior")) m.ior = (*e)["ior"].as_float();
            if (const Json* e = member(*ext, "KHR_materials_transmission")) {
                if (e->contains("transmissionFactor")) m.spec_trans = (*e)["transmissionFactor"].as_float();
                m.transmission_tex_id = tex_index(member(*e, "transmissionTexture"), toff);
            }
            if (const Json* e = member(*ext, "KHR_materials_clearcoat")) {
                if (e->contains("clearcoatFactor")) m.clearcoat = (*e)["clearcoatFactor"].as_float();
                if (e->contains("clearcoatRoughnessFactor"))
                    m.clearcoat_roughness = (*e)["clearcoatRoughnessFactor"].as_float();
                m.clearcoat_tex_id = tex_index(member(*e, "clearcoatTexture"), toff);
                m.clearcoat_roughness_tex_id = tex_index(member(*e, "clearcoatRoughnessTexture"), toff);
                m.clearcoat_normal_tex_id = tex_index(member(*e, "clearcoatNormalTexture"), toff);
            }
            if (const Json* e = member(*ext, "KHR_materials_sheen")) {
                // the reference stores sheenRoughnessFactor as the Disney sheen weight (gltf_loader.cpp:271-273)
                if (e->contains("sheenRoughnessFactor")) m.sheen = (*e)["sheenRoughnessFactor"].as_float();
                float c[3];
                if (arr3(member(*e, "sheenColorFactor"), c)) {
                    m.sheen_color_r = c[0];
                    m.sheen_color_g = c[1];
                    m.sheen_color_b = c[2];
                }
                m.sheen_color_tex_id = tex_index(member(*e, "sheenColorTexture"), toff);
                m.sheen_roughness_tex_id = tex_index(member(*e, "sheenRoughnessTexture"), toff);
            }
            if (const Json* e = member(*ext, "KHR_materials_specular")) {
                if (e->contains("specularFactor")) m.specular_factor = (*e)["specularFactor"].as_float();
                float c[3];
                if (arr3(member(*e, "specularColorFactor"), c)) {
                    m.specular_color_r = c[0];
                    m.specular_color_g = c[1];
                    m.specular_color_b = c[2];
                }
            }
            if (const Json* e = member(*ext, "KHR_materials_volume")) {
                if (e->contains("thicknessFactor")) m.thickness_factor = (*e)["thicknessFactor"].as_float();
                if (e->contains("attenuationDistance")) m.attenuation_distance = (*e)["attenuationDistance"].as_float();
                float c[3];
                if (arr3(member(*e, "attenuationColor"), c)) {
                    m.att_color_r = c[0];
                    m.att_color_g = c[1];
                    m.att_color_b = c[2];
                }
            }
            if (const Json* e = member(*ext, "KHR_materials_anisotropy")) {
                if (e->contains("anisotropyStrength")) m.anisotropic = (*e)["anisotropyStrength"].as_float();
                if (e->contains("anisotropyRotation")) m.anisotropy_rotation = (*e)["anisotropyRotation"].as_float();
            }
        }
        const std::string mode = g.contains("alphaMode") ? g["alphaMode"].str : std::string("OPAQUE");
        if (mode == "MASK") {
            m.alpha_mode = 1;
            m.alpha_cutoff = static_cast<float>(num(g, "alphaCutoff", 0.5));
        } else if (mode == "BLEND") {
            m.alpha_mode = 2;
        }
    }
}

// Node transform (gltf_loader.cpp:53-95): the column-major matrix, or T * R * S.
void node_local(const Json& n, float out[16]) {
    mat4_identity(out);
    if (const Json* m = member(n, "matrix"))
        if (m->is_array() && m->size() == 16) {
            for (int r = 0; r < 4; ++r)
                for (int c = 0; c < 4; ++c) out[r * 4 + c] = (*m)[static_cast<size_t>(c * 4 + r)].as_float();
            return;
        }
    float T[16], R[16], S[16], TR[16];
    mat4_identity(T);
    mat4_identity(R);
    mat4_identity(S);
    float t[3];
    if (arr3(member(n, "translation"), t)) {
        T[3] = t[0];
        T[7] = t[1];
        T[11] = t[2];
    }
    if (const Json* q = member(n, "rotation"))
        if (q->is_array() && q->size() == 4) {
            const float x = (*q)[0].as_float(), y = (*q)[1].as_float(), z = (*q)[2].as_float(), w = (*q)[3].as_float();
            const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, xz = x * z, yz = y * z, wx = w * x, wy = w * y,
                        wz = w * z;
            const float r[16] = {1 - 2 * (yy + zz), 2 * (xy - wz), 2 * (xz + wy), 0, 2 * (xy + wz), 1 - 2 * (xx + zz),
                                 2 * (yz - wx), 0, 2 * (xz - wy), 2 * (yz + wx), 1 - 2 * (xx + yy), 0, 0, 0, 0, 1};
            std::memcpy(R, r, sizeof(r));
        }
    float s[3];
    if (arr3(member(n, "scale"), s)) {
        S[0] = s[0];
        S[5] = s[1];
        S[10] = s[2];
    }
    mat4_mul(T, R, TR);
    mat4_mul(TR, S, out);
}

// A triangle primitive (gltf_loader.cpp:339-479).
void load_primitive(const Doc& d, const Json& prim, const float world[16], int moff, HostScene& s) {
    const int mode = inum(prim, "mode", 4);
    if (mode != 4) return;
    const Json* attrs = member(prim, "attributes");
    if (!attrs || !attrs->contains("POSITION")) return;
    const Accessor pos = accessor(d, (*attrs)["POSITION"].as_int());
    const bool has_n = attrs->contains("NORMAL"), has_uv = attrs->contains("TEXCOORD_0"),
               has_t = attrs->contains("TANGENT");
    Accessor nor, uv, tan;
    if (has_n) nor = accessor(d, (*attrs)["NORMAL"].as_int());
    if (has_uv) uv = accessor(d, (*attrs)["TEXCOORD_0"].as_int());
    if (has_t) tan = accessor(d, (*attrs)["TANGENT"].as_int());
    float inv[16], nm[16];
    if (!mat4_inverse(world, inv)) mat4_identity(inv);
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) nm[r * 4 + c] = inv[c * 4 + r];  // inverse transpose
    const uint32_t base = static_cast<uint32_t>(s.vertices.size());
    for (size_t v = 0; v < pos.count; ++v) {
        pt_vertex vert{};
        vert.position = p3(xf_point(world, V3{pos.get(v, 0), pos.get(v, 1), pos.get(v, 2)}));
        if (has_n && v < nor.count) vert.normal = p3(normalized(xf_vector(nm, V3{nor.get(v, 0), nor.get(v, 1), nor.get(v, 2)})));
        if (has_uv && v < uv.count) {
            vert.u = uv.get(v, 0);
            vert.v = uv.get(v, 1);
        }
        if (has_t && v < tan.count) {
            const V3 t = normalized(xf_vector(world, V3{tan.get(v, 0), tan.get(v, 1), tan.get(v, 2)}));
            vert.tangent[0] = t.x;
            vert.tangent[1] = t.y;
            vert.tangent[2] = t.z;
            vert.tangent[3] = tan.ncomp >= 4 ? tan.get(v, 3) : 1.0f;
        }
        s.vertices.push_back(vert);
    }
    const int mat = prim.contains("material") ? prim["material"].as_int() + moff : 0;
    const size_t face0 = s.faces.size();
    if (prim.contains("indices")) {
        const Accessor ix = accessor(d, prim["indices"].as_int());
        if (ix.ctype != 5121 && ix.ctype != 5123 && ix.ctype != 5125) throw std::runtime_error("glTF: bad index type");
        for (size_t t = 0; t + 2 < ix.count; t += 3) {
            const uint32_t i0 = ix.index(t), i1 = ix.index(t + 1), i2 = ix.index(t + 2);
            if (i0 >= pos.count || i1 >= pos.count || i2 >= pos.count) throw std::runtime_error("glTF: index out of range");
            s.faces.push_back(pt_face{base + i0, base + i1, base + i2, mat});
        }
    } else {
        for (size_t t = 0; t + 2 < pos.count; t += 3)
            s.faces.push_back(pt_face{base + static_cast<uint32_t>(t), base + static_cast<uint32_t>(t + 1),
                                      base + static_cast<uint32_t>(t + 2), mat});
    }
    if (!has_n) {  // area-weighted face normals (gltf_loader.cpp:455-478)
        for (size_t v = 0; v < pos.count; ++v) s.vertices[base + v].normal = pt_vec3p{0, 0, 0, 0};
        for (size_t f = face0; f < s.faces.size(); ++f) {
            const pt_face& F = s.faces[f];
            const pt_vec3p &a = s.vertices[F.v0].position, &b = s.vertices[F.v1].position, &c = s.vertices[F.v2].position;
            const V3 e1{b.x - a.x, b.y - a.y, b.z - a.z}, e2{c.x - a.x, c.y - a.y, c.z - a.z};
            const V3 n{e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x};
            for (uint32_t vi : {F.v0, F.v1, F.v2}) {
                s.vertices[vi].normal.x += n.x;
                s.vertices[vi].normal.y += n.y;
                s.vertices[vi].normal.z += n.z;
            }
        }
        for (size_t v = 0; v < pos.count; ++v) {
            pt_vec3p& n = s.vertices[base + v].normal;
            const float l = std::sqrt(n.x * n.x + n.y * n.y + n.z * n.z);
            if (l > 0.0f) {
                n.x /= l;
                n.y /= l;
                n.z /= l;
            }
        }
    }
}

struct LightDef {
    int type = 0;  // point
    float color[3] = {1, 1, 1};
    float intensity = 1.0f, range = 1e30f, inner = 0.0f, outer = 0.7853982f;
};

std::vector<LightDef> light_defs(const Doc& d) {
    std::vector<LightDef> out;
    const Json* ext = member(d.j, "extensions");
    const Json* kl = ext ? member(*ext, "KHR_lights_punctual") : nullptr;
    const Json* ls = kl ? member(*kl, "lights") : nullptr;
    if (!ls || !ls->is_array()) return out;
    for (const Json& l : ls->arr) {
        LightDef def;
        const std::string t = l.contains("type") ? l["type"].str : std::string("point");
        def.type = t == "directional" ? 1 : (t == "spot" ? 2 : 0);
        arr3(member(l, "color"), def.color);
        def.intensity = static_cast<float>(num(l, "intensity", 1.0));
        def.range = static_cast<float>(num(l, "range", 1e30));
        if (const Json* sp = member(l, "spot")) {
            def.inner = static_cast<float>(num(*sp, "innerConeAngle", 0.0));
            def.outer = static_cast<float>(num(*sp, "outerConeAngle", 0.7853982));
        }
        out.push_back(def);
    }
    return out;
}

void visit(const Doc& d, int ni, const float parent[16], int moff, const std::vector<LightDef>& lights, HostScene& s,
           bool& camera_set, int depth) {
    const Json& nodes = d.list("nodes");
    if (ni < 0 || static_cast<size_t>(ni) >= nodes.size() || depth > 256) return;
    const Json& n = nodes[static_cast<size_t>(ni)];
    float local[16], world[16];
    node_local(n, local);
    mat4_mul(parent, local, world);
    const Json& meshes = d.list("meshes");
    const int mi = inum(n, "mesh", -1);
    if (mi >= 0 && static_cast<size_t>(mi) < meshes.size())
        if (const Json* prims = member(meshes[static_cast<size_t>(mi)], "primitives"))
            for (const Json& p : prims->arr) load_primitive(d, p, world, moff, s);
    const Json& cams = d.list("cameras");
    const int ci = inum(n, "camera", -1);
    if (!camera_set && ci >= 0 && static_cast<size_t>(ci) < cams.size()) {
        const Json& c = cams[static_cast<size_t>(ci)];
        if (c.contains("type") && c["type"].str == "perspective" && c.contains("perspective")) {
            const Json& p = c["perspective"];
            const V3 pos = xf_point(world, V3{0, 0, 0});
            const V3 fwd = normalized(xf_vector(world, V3{0, 0, -1}));
            s.camera.position = p3(pos);
            s.camera.look_at = pt_vec3p{pos.x + fwd.x, pos.y + fwd.y, pos.z + fwd.z, 0.0f};
            s.camera.up = p3(normalized(xf_vector(world, V3{0, 1, 0})));
            s.camera.fov = static_cast<float>(num(p, "yfov", 0.0) * 180.0 / 3.14159265358979323846);
            const double aspect = num(p, "aspectRatio", 0.0);
            if (aspect > 0.0) s.camera.width = static_cast<int>(aspect * s.camera.height);
            camera_set = true;
        }
    }
    if (const Json* ext = member(n, "extensions"))
        if (const Json* kl = member(*ext, "KHR_lights_punctual"))
            if (kl->contains("light")) {
                const int li = (*kl)["light"].as_int();
                if (li >= 0 && static_cast<size_t>(li) < lights.size()) {
                    const LightDef& def = lights[static_cast<size_t>(li)];
                    pt_punctual_light pl = default_light();
                    pl.type = def.type;
                    pl.color = pt_vec3p{def.color[0], def.color[1], def.color[2], 0.0f};
                    pl.intensity = def.intensity;
                    pl.range = def.range;
                    pl.position = p3(xf_point(world, V3{0, 0, 0}));
                    pl.direction = p3(normalized(xf_vector(world, V3{0, 0, -1})));
                    pl.inner_cone = std::cos(def.inner);
                    pl.outer_cone = std::cos(def.outer);
                    s.lights.push_back(pl);
                }
            }
    if (const Json* ch = member(n, "children"))
        for (const Json& c : ch->arr) visit(d, c.as_int(), world, moff, lights, s, camera_set, depth + 1);
}

bool parse_document(const std::string& path, Doc& d, std::string* err) {
    std::vector<uint8_t> file;
    if (!read_file(path, file, err)) return false;
    const size_t slash = path.find_last_of('/');
    d.dir = slash == std::string::npos ? std::string() : path.substr(0, slash + 1);
    std::string text;
    std::vector<uint8_t> bin;
    bool glb = false;
    auto le32 = [&](size_t o) {
        return static_cast<uint32_t>(file[o]) | (static_cast<uint32_t>(file[o + 1]) << 8) |
               (static_cast<uint32_t>(file[o + 2]) << 16) | (static_cast<uint32_t>(file[o + 3]) << 24);
    };
    if (file.size() >= 12 && le32(0) == 0x46546C67u) {  // "glTF": binary container
        glb = true;
        if (le32(4) != 2) {
            if (err) *err = "GLB: only version 2 is supported";
            return false;
        }
        size_t o = 12;
        const size_t total = std::min<size_t>(le32(8), file.size());
        while (o + 8 <= total) {
            const uint32_t len = le32(o), type = le32(o + 4);
            if (o + 8 + len > total) {
                if (err) *err = "GLB: truncated chunk";
                return false;
            }
            if (type == 0x4E4F534Au && text.empty()) text.assign(reinterpret_cast<const char*>(&file[o + 8]), len);
            else if (type == 0x004E4942u && bin.empty()) bin.assign(file.begin() + static_cast<std::ptrdiff_t>(o + 8),
                                                                    file.begin() + static_cast<std::ptrdiff_t>(o + 8 + len));
            o += 8 + ((len + 3) & ~3u);
        }
    } else {
        text.assign(file.begin(), file.end());
    }
    std::string perr;
    if (!json_parse(text, d.j, &perr)) {
        if (err) *err = "glTF JSON: " + perr;
        return false;
    }
    const Json& bufs = d.list("buffers");
    d.buffers.resize(bufs.size());
    for (size_t i = 0; i < bufs.size(); ++i) {
        if (bufs[i].contains("uri")) {
            if (!load_uri(bufs[i]["uri"].str, d.dir, d.buffers[i], err)) return false;
        } else if (glb && i == 0) {
            d.buffers[i] = bin;
        } else {
            if (err) *err = "glTF: buffer " + std::to_string(i) + " has no data";
            return false;
        }
    }
    return true;
}

}  // namespace

bool load_gltf(const std::string& path, HostScene& s, const GltfOptions& opts, std::string* err) {
    Doc d;
    try {
        if (!parse_document(path, d, err)) return false;
        const int toff = static_cast<int>(s.textures.size());
        const int moff = static_cast<int>(s.materials.size());
        std::vector<Image> texs;
        load_textures(d, texs, opts.warnings);
        for (auto& t : texs) s.textures.push_back(std::move(t));
        std::vector<pt_material> mats;
        load_materials(d, mats, toff, opts.material_names);
        s.materials.insert(s.materials.end(), mats.begin(), mats.end());
        if (s.materials.empty()) s.materials.push_back(default_material());
        const std::vector<LightDef> lights = light_defs(d);
        bool camera_set = !opts.load_camera;
        const Json& scenes = d.list("scenes");
        int si = inum(d.j, "scene", -1);
        if (si < 0) si = 0;
        if (static_cast<size_t>(si) < scenes.size())
            if (const Json* roots = member(scenes[static_cast<size_t>(si)], "nodes"))
                for (const Json& r : roots->arr) visit(d, r.as_int(), opts.transform, moff, lights, s, camera_set, 0);
        if (opts.auto_fit_camera && !camera_set && !s.vertices.empty()) {  // gltf_loader.cpp:650-679
            float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
            for (const auto& v : s.vertices) {
                lo[0] = std::fmin(lo[0], v.position.x);
                lo[1] = std::fmin(lo[1], v.position.y);
                lo[2] = std::fmin(lo[2], v.position.z);
                hi[0] = std::fmax(hi[0], v.position.x);
                hi[1] = std::fmax(hi[1], v.position.y);
                hi[2] = std::fmax(hi[2], v.position.z);
            }
            const V3 c{(lo[0] + hi[0]) * 0.5f, (lo[1] + hi[1]) * 0.5f, (lo[2] + hi[2]) * 0.5f};
            const V3 dg{hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
            float radius = std::sqrt(dg.x * dg.x + dg.y * dg.y + dg.z * dg.z) * 0.5f;
            if (radius < 1e-5f) radius = 1.0f;
            const float fov = s.camera.fov * 3.14159265358979323846f / 180.0f;
            const float dist = radius / std::tan(fov * 0.5f) * 1.2f;
            s.camera.position = pt_vec3p{c.x, c.y + radius * 0.3f, c.z + dist, 0.0f};
            s.camera.look_at = p3(c);
            s.camera.up = pt_vec3p{0, 1, 0, 0};
        }
        if (opts.load_environment && s.env.empty()) {  // gltf_loader.cpp:681-697
            namespace fs = std::filesystem;
            std::error_code ec;
            fs::path dir = fs::absolute(fs::path(path), ec).parent_path();
            while (!dir.empty()) {
                const fs::path cand = dir / "models" / "kloofendal_48d_partly_cloudy_puresky_4k.hdr";
                if (fs::exists(cand, ec)) {
                    std::string herr;
                    load_image(cand.string(), s.env, &herr);
                    break;
                }
                if (dir == dir.parent_path()) break;
                dir = dir.parent_path();
            }
        }
        const bool has_env = !s.env.empty();
        bool has_emissive = false;
        for (const auto& m : s.materials) has_emissive = has_emissive || material_is_emissive(m);
        if (opts.add_default_lights && !has_env && !has_emissive && s.lights.empty()) {  // :699-722
            pt_punctual_light key = default_light(), fill = default_light();
            key.type = fill.type = 1;
            key.direction = p3(normalized(V3{-0.5f, -0.8f, -0.6f}));
            key.color = pt_vec3p{1.0f, 0.95f, 0.9f, 0.0f};
            key.intensity = 3.0f;
            fill.direction = p3(normalized(V3{0.6f, -0.3f, 0.5f}));
            fill.color = pt_vec3p{0.7f, 0.8f, 1.0f, 0.0f};
            fill.intensity = 1.0f;
            s.lights.push_back(key);
            s.lights.push_back(fill);
        }
    } catch (const std::exception& e) {
        if (err) *err = std::string("'") + path + "': " + e.what();
        return false;
    }
    return true;
}

}  // namespace ptamd

// Light-selection tables built on the host once per scene:
//  * a Vose alias table over emissive triangles weighted by area * luminance
//    of the emitted radiance, and a per-face selection pmf;
//  * a Vose alias table over environment texels weighted by luminance *
//    sin(theta) of the texel centre, and the per-texel pmf.
// Semantics follow Part_2/src/lights/light_sampler.cpp:7-149 exactly (single
// precision sums in texel order, FIFO small/large work lists, leftovers set to
// probability 1), so the tables are bit-identical to the reference's.
#include <cmath>
#include <deque>

#include "host_scene.h"

namespace ptamd {

namespace {
inline float lum(float r, float g, float b) { return 0.2126f * r + 0.7152f * g + 0.0722f * b; }
}  // namespace

// light_sampler.cpp:7-55
std::vector<AliasEntry> build_alias_table(const std::vector<float>& weights) {
    const uint32_t n = static_cast<uint32_t>(weights.size());
    if (n == 0) return {};
    float sum = 0.0f;
    for (float w : weights) sum += w;
    if (sum <= 0.0f) return {};
    std::vector<float> prob(n);
    for (uint32_t i = 0; i < n; ++i) prob[i] = weights[i] / sum * static_cast<float>(n);
    std::vector<AliasEntry> table(n);
    std::deque<uint32_t> small, large;
    for (uint32_t i = 0; i < n; ++i) (prob[i] < 1.0f ? small : large).push_back(i);
    while (!small.empty() && !large.empty()) {
        uint32_t s = small.front(); small.pop_front();
        uint32_t l = large.front(); large.pop_front();
        table[s].prob = prob[s];
        table[s].alias = l;
        prob[l] = (prob[l] + prob[s]) - 1.0f;
        (prob[l] < 1.0f ? small : large).push_back(l);
    }
    while (!large.empty()) {
        uint32_t l = large.front(); large.pop_front();
        table[l] = {1.0f, l};
    }
    while (!small.empty()) {
        uint32_t s = small.front(); small.pop_front();
        table[s] = {1.0f, s};
    }
    return table;
}

static float face_area(const pt_vertex& a, const pt_vertex& b, const pt_vertex& c) {
    float e1x = b.position.x - a.position.x, e1y = b.position.y - a.position.y, e1z = b.position.z - a.position.z;
    float e2x = c.position.x - a.position.x, e2y = c.position.y - a.position.y, e2z = c.position.z - a.position.z;
    float cx = e1y * e2z - e1z * e2y, cy = e1z * e2x - e1x * e2z, cz = e1x * e2y - e1y * e2x;
    return 0.5f * std::sqrt(cx * cx + cy * cy + cz * cz);
}

// light_sampler.cpp:57-110
void build_emissive_tables(const std::vector<pt_vertex>& v, const std::vector<pt_face>& f,
                           const std::vector<pt_material>& m, LightTables& out) {
    out.emissive.clear();
    out.tri_pmf.assign(f.size(), 0.0f);
    out.tri_alias.clear();
    std::vector<float> weights;
    for (uint32_t i = 0; i < static_cast<uint32_t>(f.size()); ++i) {
        const pt_face& face = f[i];
        if (face.material_id < 0 || face.material_id >= static_cast<int>(m.size())) continue;
        const pt_material& mat = m[face.material_id];
        if (!material_is_emissive(mat)) continue;
        float area = face_area(v[face.v0], v[face.v1], v[face.v2]);
        if (area <= 0.0f) continue;
        float le = lum(mat.emissive.x * mat.emissive_strength, mat.emissive.y * mat.emissive_strength,
                       mat.emissive.z * mat.emissive_strength);
        float w = area * std::fmax(le, 0.0f);
        if (w <= 0.0f) continue;
        out.emissive.push_back({i, 0.0f, 0.0f, 0.0f});
        weights.push_back(w);
    }
    if (out.emissive.empty()) return;
    float sum = 0.0f;
    for (float w : weights) sum += w;
    if (sum <= 0.0f) {
        out.emissive.clear();
        return;
    }
    float accum = 0.0f;
    for (size_t i = 0; i < out.emissive.size(); ++i) {
        float pmf = weights[i] / sum;
        accum += pmf;
        out.emissive[i].pmf = pmf;
        out.emissive[i].cdf = accum;
        out.tri_pmf[out.emissive[i].tri] = pmf;
    }
    out.emissive.back().cdf = 1.0f;
    out.tri_alias = build_alias_table(weights);
}

// light_sampler.cpp:112-149
void build_env_tables(const Image& env, LightTables& out) {
    out.env_pmf.clear();
    out.env_alias.clear();
    out.env_w = out.env_h = 0;
    if (env.width <= 0 || env.height <= 0 || env.rgba.empty()) return;
    const int n = env.width * env.height;
    std::vector<float> weights(static_cast<size_t>(n), 0.0f);
    const float pi = 3.14159265358979323846f;
    for (int y = 0; y < env.height; ++y) {
        float theta = pi * (static_cast<float>(y) + 0.5f) / static_cast<float>(env.height);
        float st = std::fmax(std::sin(theta), 0.0f);
        for (int x = 0; x < env.width; ++x) {
            size_t idx = static_cast<size_t>(y) * env.width + x;
            const float* c = &env.rgba[idx * 4];
            weights[idx] = std::fmax(lum(c[0], c[1], c[2]), 0.0f) * st;
        }
    }
    float sum = 0.0f;
    for (float w : weights) sum += w;
    if (sum <= 0.0f) return;
    out.env_w = env.width;
    out.env_h = env.height;
    out.env_pmf.resize(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) out.env_pmf[i] = weights[i] / sum;
    out.env_alias = build_alias_table(weights);
}

}  // namespace ptamd

// Oracle: CPU restatement of the reference renderer's per-pixel integrator.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).  Parity pins of every function are
// listed in oracle/README.md; functions without a pin say "parity unpinned"
// there.  Paths below are relative to Part_2/src of the reference.
//
// Structure mirrors the reference one file at a time: samplers/, accel/,
// geometry/, materials/, lights/, integrators/.  Everything that the
// integrator computes is templated on the real type R (float = the
// reference's arithmetic; double = the converged FP64 oracle).  Float builds
// are compiled with -ffp-contract=off so every expression rounds as written.
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <deque>
#include <thread>
#include <vector>

namespace orc {

// The math functions the integrator's templates call: std's for float and
// double; the reverse-mode type of the host adjoint (below) overloads them.
template <class R> inline R msqrt(R x) { return std::sqrt(x); }
template <class R> inline R mpow(R a, R b) { return std::pow(a, b); }
template <class R> inline R mfabs(R x) { return std::fabs(x); }

// A parameter read (material factor, texel): its value; the host adjoint's
// reverse-mode type turns reads of registered parameter arrays into leaves.
template <class R> inline R paramRead(const float* p) { return R(*p); }

// ------------------------------------------------------------- math/vecmath.h
template <class R>
struct V3 {
    R x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(R a, R b, R c) : x(a), y(b), z(c) {}
    explicit V3(R s) : x(s), y(s), z(s) {}
    V3 operator+(const V3& b) const { return {x + b.x, y + b.y, z + b.z}; }
    V3 operator-(const V3& b) const { return {x - b.x, y - b.y, z - b.z}; }
    V3 operator*(R s) const { return {x * s, y * s, z * s}; }
    V3 operator*(const V3& b) const { return {x * b.x, y * b.y, z * b.z}; }
    V3 operator/(R s) const { return {x / s, y / s, z / s}; }
    V3 operator-() const { return {-x, -y, -z}; }
    V3& operator+=(const V3& b) { x += b.x; y += b.y; z += b.z; return *this; }
    V3& operator*=(const V3& b) { x *= b.x; y *= b.y; z *= b.z; return *this; }
    V3& operator/=(R s) { x /= s; y /= s; z /= s; return *this; }
    R operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    bool operator!=(const V3& b) const { return x != b.x || y != b.y || z != b.z; }
};
template <class R> V3<R> operator*(R s, const V3<R>& v) { return v * s; }
template <class R> R dot(const V3<R>& a, const V3<R>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> V3<R> cross(const V3<R>& a, const V3<R>& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> R length(const V3<R>& a) { return msqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
template <class R> V3<R> normalize(const V3<R>& a) {  // vecmath.h:96
    R l = length(a);
    return l > 0 ? a / l : V3<R>();
}
template <class R> V3<R> reflect(const V3<R>& i, const V3<R>& n) { return i - R(2) * dot(i, n) * n; }
template <class R> R fmn(R a, R b) { return std::fmin(a, b); }  // math_utils.h:13-27 (NaN dropping)
template <class R> R fmx(R a, R b) { return std::fmax(a, b); }
template <class R> R clampr(R x, R lo, R hi) { return fmn(fmx(x, lo), hi); }
template <class R> V3<R> v3of(const pt_vec3p& p) { return V3<R>(R(p.x), R(p.y), R(p.z)); }

template <class R> struct K {
    static constexpr R Pi = R(3.14159265358979323846);
    static constexpr R InvPi = R(0.31830988618379067154);
    static constexpr R Inv2Pi = R(0.15915494309189533577);
    static constexpr R TwoPi = R(6.28318530717958647692);
};

// ---------------------------------------------------------------- math/frame.h:11-34
template <class R>
struct Frame {
    V3<R> t, b, n;
    explicit Frame(const V3<R>& nn) : n(nn) {
        V3<R> helper(1, 0, 0);
        if (mfabs(nn.x) > R(0.999)) helper = V3<R>(0, 0, 1);
        t = normalize(cross(nn, helper));
        b = normalize(cross(nn, t));
    }
    V3<R> toWorld(const V3<R>& v) const { return v.x * t + v.y * b + v.z * n; }
    V3<R> toLocal(const V3<R>& v) const { return V3<R>(dot(v, t), dot(v, b), dot(v, n)); }
};

// ------------------------------------------------------------- math/sampling.h
template <class R> V3<R> sampleCosineHemisphere(R u1, R u2) {  // :18-25
    R r = msqrt(u1);
    R phi = K<R>::TwoPi * u2;
    return V3<R>(r * std::cos(phi), r * std::sin(phi), msqrt(fmx(R(0), R(1) - u1)));
}
template <class R> R powerHeuristic(R a, R b) {  // :34-39
    R a2 = a * a, b2 = b * b, s = a2 + b2;
    return s > 0 ? a2 / s : R(0);
}

// ------------------------------------------------------------------ samplers/
// sobol.cuh:14-24 (host branch)
inline uint32_t reverseBits32(uint32_t x) {
    x = ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
    x = ((x >> 2) & 0x33333333u) | ((x & 0x33333333u) << 2);
    x = ((x >> 4) & 0x0F0F0F0Fu) | ((x & 0x0F0F0F0Fu) << 4);
    x = ((x >> 8) & 0x00FF00FFu) | ((x & 0x00FF00FFu) << 8);
    return (x >> 16) | (x << 16);
}
inline uint32_t hashMix(uint32_t x) {  // sobol.cuh:48-57
    x ^= x >> 17; x ^= x * 0xed5ad4bbu; x ^= x >> 11; x ^= x * 0xac4c1b51u;
    x ^= x >> 15; x ^= x * 0x31848babu; x ^= x >> 14;
    return x;
}
inline uint32_t owenHash(uint32_t x, uint32_t seed) {  // :59-66
    x += seed; x ^= x * 0x6c50b47cu; x ^= x * 0xb82f1e52u; x ^= x * 0xc7afe638u; x ^= x * 0x8d22f6e6u;
    return x;
}
inline uint32_t owenScramble(uint32_t v, uint32_t seed) {  // :69-74
    return reverseBits32(owenHash(reverseBits32(v), seed));
}
// :77-85, plus the build's global seed (0 = reference)
inline uint32_t dimensionSeed(int px, int py, uint32_t dim, uint32_t gseed) {
    uint32_t h = static_cast<uint32_t>(px);
    h = hashMix(h + 0x9e3779b9u);
    h ^= static_cast<uint32_t>(py) + 0x9e3779b9u + (h << 6) + (h >> 2);
    h = hashMix(h);
    if (gseed != 0u) h ^= hashMix(gseed ^ 0x85ebca6bu);
    h ^= dim + 0x9e3779b9u + (h << 6) + (h >> 2);
    h = hashMix(h);
    return h;
}
inline uint32_t sobolU32(int px, int py, uint32_t index, uint32_t dim, uint32_t gseed) {  // :107-117
    uint32_t hash = dimensionSeed(px, py, dim, gseed);
    uint32_t shuffled = owenScramble(index, hash);
    uint32_t raw = reverseBits32(shuffled);  // sobolDim0
    return owenScramble(raw, hashMix(hash));
}
inline float sobolFloat(uint32_t s) { return std::fmin(static_cast<float>(s) * (1.0f / 4294967296.0f), 0.99999994f); }

// samplers/pcg.cuh + sampler.cuh
struct Sampler {
    int type = 0;
    uint32_t state = 0;
    uint32_t index = 0, dim = 0;
    int px = 0, py = 0;
    uint32_t gseed = 0;
    void initPCG(int x, int y, uint32_t frame, uint32_t sample, uint32_t seed) {  // pcg.cuh:18-27
        type = 0;
        uint32_t s = static_cast<uint32_t>(x) * 1973u + static_cast<uint32_t>(y) * 9277u + frame * 26699u +
                     sample * 12149u + seed * 0x9E3779B9u;
        s = s * 747796405u + 2891336453u;
        state = s | 1u;
    }
    void initSobol(int x, int y, uint32_t idx, uint32_t seed) {  // sobol.cuh:99-104
        type = 1; px = x; py = y; index = idx; dim = 0; gseed = seed;
    }
    uint32_t nextU32() {
        if (type == 1) return sobolU32(px, py, index, dim++, gseed);
        uint32_t old = state;  // pcg.cuh:11-16
        state = old * 747796405u + 2891336453u;
        uint32_t word = ((old >> ((old >> 28u) + 4u)) ^ old) * 277803737u;
        return (word >> 22u) ^ word;
    }
    float get1D() {  // sampler.cuh:25-29
        uint32_t u = nextU32();
        if (type == 1) return sobolFloat(u);
        return static_cast<float>(u) / 4294967296.0f;
    }
};

// ---------------------------------------------------------- accel/bvh_common.h
struct Node {
    float bmin[3];
    uint32_t data0;
    float bmax[3];
    uint32_t data1;
    bool isLeaf() const { return (data0 & 0x80000000u) != 0; }
    uint32_t first() const { return data0 & 0x7FFFFFFFu; }
    int maxDim() const {
        float dx = bmax[0] - bmin[0], dy = bmax[1] - bmin[1], dz = bmax[2] - bmin[2];
        if (dx > dy && dx > dz) return 0;
        return dy > dz ? 1 : 2;
    }
    float centroid(int a) const { return 0.5f * (bmin[a] + bmax[a]); }
};

struct Prim {
    float bmin[3], bmax[3];
    uint32_t index;
    float centroid(int a) const { return 0.5f * (bmin[a] + bmax[a]); }
};

// ------------------------------------------------------- accel/bvh_builder.cpp
class Builder {
public:
    int buckets = 24, leafThreshold = 4;
    std::vector<Node> nodes;
    std::vector<Prim> prims;

    void build(const std::vector<pt_vertex>& v, std::vector<pt_face>& faces, std::vector<uint32_t>& order) {
        prims.resize(faces.size());
        for (size_t i = 0; i < faces.size(); ++i) {  // :62-75
            const pt_vertex &a = v[faces[i].v0], &b = v[faces[i].v1], &c = v[faces[i].v2];
            Prim& p = prims[i];
            p.index = static_cast<uint32_t>(i);
            p.bmin[0] = std::fmin(std::fmin(a.position.x, b.position.x), c.position.x);
            p.bmin[1] = std::fmin(std::fmin(a.position.y, b.position.y), c.position.y);
            p.bmin[2] = std::fmin(std::fmin(a.position.z, b.position.z), c.position.z);
            p.bmax[0] = std::fmax(std::fmax(a.position.x, b.position.x), c.position.x);
            p.bmax[1] = std::fmax(std::fmax(a.position.y, b.position.y), c.position.y);
            p.bmax[2] = std::fmax(std::fmax(a.position.z, b.position.z), c.position.z);
        }
        raw.clear();
        raw.reserve(2 * faces.size());
        if (!faces.empty()) buildNode(0, static_cast<uint32_t>(faces.size()));
        // :179-194 depth-first reorder (left child = next node)
        nodes.clear();
        nodes.reserve(raw.size());
        if (!raw.empty()) reorder(0);
        // :196-201 faces permuted in place
        std::vector<pt_face> re(faces.size());
        order.resize(faces.size());
        for (size_t i = 0; i < faces.size(); ++i) {
            re[i] = faces[prims[i].index];
            order[i] = prims[i].index;
        }
        faces.swap(re);
    }

private:
    std::vector<Node> raw;

    static void reset(float lo[3], float hi[3]) {
        lo[0] = lo[1] = lo[2] = FLT_MAX;
        hi[0] = hi[1] = hi[2] = -FLT_MAX;
    }
    static float sa(const float lo[3], const float hi[3]) {  // :35-41
        float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (dx < 0 || dy < 0 || dz < 0) return 0.0f;
        return 2.0f * (dx * dy + dx * dz + dy * dz);
    }
    static int maxDim(const float lo[3], const float hi[3]) {  // :43-49
        float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (dx > dy && dx > dz) return 0;
        return dy > dz ? 1 : 2;
    }
    uint32_t buildNode(uint32_t start, uint32_t end) {  // :92-122
        uint32_t idx = static_cast<uint32_t>(raw.size());
        raw.push_back(Node{});
        float lo[3], hi[3];
        reset(lo, hi);
        for (uint32_t i = start; i < end; ++i)
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::fmin(lo[k], prims[i].bmin[k]);
                hi[k] = std::fmax(hi[k], prims[i].bmax[k]);
            }
        uint32_t count = end - start;
        auto leaf = [&]() {
            Node n;
            for (int k = 0; k < 3; ++k) { n.bmin[k] = lo[k]; n.bmax[k] = hi[k]; }
            n.data0 = start | 0x80000000u;
            n.data1 = count;
            raw[idx] = n;
            return idx;
        };
        if (static_cast<int>(count) < leafThreshold) return leaf();
        uint32_t mid = sah(start, end, lo, hi);
        if (mid == end) return leaf();
        uint32_t l = buildNode(start, mid);
        uint32_t r = buildNode(mid, end);
        Node n;
        for (int k = 0; k < 3; ++k) { n.bmin[k] = lo[k]; n.bmax[k] = hi[k]; }
        n.data0 = l;
        n.data1 = r;
        raw[idx] = n;
        return idx;
    }
    uint32_t sah(uint32_t start, uint32_t end, const float lo[3], const float hi[3]) {  // :124-177
        float cmin[3], cmax[3];
        reset(cmin, cmax);
        for (uint32_t i = start; i < end; ++i)
            for (int k = 0; k < 3; ++k) {
                float c = prims[i].centroid(k);
                cmin[k] = std::fmin(cmin[k], c);
                cmax[k] = std::fmax(cmax[k], c);
            }
        int dim = maxDim(cmin, cmax);
        const int nb = buckets, ns = nb - 1;
        std::vector<float> bmin(3 * nb, FLT_MAX), bmax(3 * nb, -FLT_MAX);
        std::vector<uint32_t> cnt(nb, 0);
        auto bucketOf = [&](const Prim& p) {
            float c = p.centroid(dim);
            float d = cmax[dim] - cmin[dim];
            float off = d > 0 ? (c - cmin[dim]) / d : 0.0f;
            int b = static_cast<int>(nb * off);
            return b >= nb ? nb - 1 : b;
        };
        for (uint32_t i = start; i < end; ++i) {
            int b = bucketOf(prims[i]);
            cnt[b]++;
            for (int k = 0; k < 3; ++k) {
                bmin[3 * b + k] = std::fmin(bmin[3 * b + k], prims[i].bmin[k]);
                bmax[3 * b + k] = std::fmax(bmax[3 * b + k], prims[i].bmax[k]);
            }
        }
        std::vector<float> costs(ns, 0.0f);
        uint32_t cf = 0;
        float flo[3], fhi[3];
        reset(flo, fhi);
        for (int i = 0; i < ns; ++i) {
            cf += cnt[i];
            for (int k = 0; k < 3; ++k) {
                flo[k] = std::fmin(flo[k], bmin[3 * i + k]);
                fhi[k] = std::fmax(fhi[k], bmax[3 * i + k]);
            }
            costs[i] += cf * sa(flo, fhi);
        }
        uint32_t cbk = 0;
        float blo[3], bhi[3];
        reset(blo, bhi);
        for (int i = ns - 1; i >= 0; --i) {
            cbk += cnt[i + 1];
            for (int k = 0; k < 3; ++k) {
                blo[k] = std::fmin(blo[k], bmin[3 * (i + 1) + k]);
                bhi[k] = std::fmax(bhi[k], bmax[3 * (i + 1) + k]);
            }
            costs[i] += cbk * sa(blo, bhi);
        }
        auto it = std::min_element(costs.begin(), costs.end());
        float minCost = *it;
        int minBucket = static_cast<int>(it - costs.begin());
        float leafCost = static_cast<float>(end - start);
        float lsa = sa(lo, hi);
        minCost = 0.5f + (lsa > 0.0f ? minCost / lsa : 0.0f);
        if (leafCost < minCost) return end;
        auto m = std::partition(prims.begin() + start, prims.begin() + end,
                                [&](const Prim& p) { return bucketOf(p) <= minBucket; });
        return static_cast<uint32_t>(m - prims.begin());
    }
    uint32_t reorder(uint32_t old) {
        nodes.push_back(raw[old]);
        uint32_t cur = static_cast<uint32_t>(nodes.size() - 1);
        if (!raw[old].isLeaf()) {
            uint32_t l = reorder(raw[old].data0);
            nodes[cur].data0 = l;
            uint32_t r = reorder(raw[old].data1);
            nodes[cur].data1 = r;
        }
        return cur;
    }
};

// --------------------------------------------------------------- scene data
struct Texture {
    int w = 0, h = 0, wrapS = 0, wrapT = 0, linear = 0;
    std::vector<float> rgba;
};

struct AliasEntry {
    float prob;
    uint32_t alias;
};

}  // namespace orc

struct oracle_scene {
    std::vector<pt_vertex> v;
    std::vector<pt_face> f;  // BVH order
    std::vector<uint32_t> order;
    std::vector<pt_material> m;
    std::vector<orc::Node> nodes;
    std::vector<pt_punctual_light> lights;
    std::vector<orc::Texture> tex;
    orc::Texture env;
    // light_sampler.h:36-48
    struct Emissive { uint32_t tri; float pmf, cdf; };
    std::vector<Emissive> emissive;
    std::vector<float> triPmf;
    std::vector<orc::AliasEntry> triAlias;
    std::vector<float> envPmf;
    std::vector<orc::AliasEntry> envAlias;
};

namespace orc {

// ---------------------------------------------------- lights/light_sampler.cpp
std::vector<AliasEntry> buildAliasTable(const std::vector<float>& w) {  // :7-55
    const uint32_t n = static_cast<uint32_t>(w.size());
    if (n == 0) return {};
    float sum = 0.0f;
    for (float x : w) sum += x;
    if (sum <= 0.0f) return {};
    std::vector<float> prob(n);
    for (uint32_t i = 0; i < n; ++i) prob[i] = w[i] / sum * static_cast<float>(n);
    std::vector<AliasEntry> t(n);
    std::deque<uint32_t> small, large;
    for (uint32_t i = 0; i < n; ++i) {
        if (prob[i] < 1.0f) small.push_back(i);
        else large.push_back(i);
    }
    while (!small.empty() && !large.empty()) {
        uint32_t s = small.front(); small.pop_front();
        uint32_t l = large.front(); large.pop_front();
        t[s].prob = prob[s];
        t[s].alias = l;
        prob[l] = (prob[l] + prob[s]) - 1.0f;
        if (prob[l] < 1.0f) small.push_back(l);
        else large.push_back(l);
    }
    while (!large.empty()) { uint32_t l = large.front(); large.pop_front(); t[l] = {1.0f, l}; }
    while (!small.empty()) { uint32_t s = small.front(); small.pop_front(); t[s] = {1.0f, s}; }
    return t;
}

inline float luminance(float r, float g, float b) { return 0.2126f * r + 0.7152f * g + 0.0722f * b; }

inline bool isEmissive(const pt_material& m) {  // material.h:83-86
    return m.emissive_strength > 0.0f && (m.emissive.x > 0.0f || m.emissive.y > 0.0f || m.emissive.z > 0.0f);
}

inline float faceArea(const pt_vertex& a, const pt_vertex& b, const pt_vertex& c) {  // light_sampler.h:54-56
    V3<float> e1 = v3of<float>(b.position) - v3of<float>(a.position);
    V3<float> e2 = v3of<float>(c.position) - v3of<float>(a.position);
    return 0.5f * length(cross(e1, e2));
}

void buildLightTables(oracle_scene& s) {
    // :57-110
    s.emissive.clear();
    s.triPmf.assign(s.f.size(), 0.0f);
    std::vector<float> w;
    for (uint32_t i = 0; i < s.f.size(); ++i) {
        const pt_face& fc = s.f[i];
        if (fc.material_id < 0 || fc.material_id >= static_cast<int>(s.m.size())) continue;
        const pt_material& mt = s.m[fc.material_id];
        if (!isEmissive(mt)) continue;
        float area = faceArea(s.v[fc.v0], s.v[fc.v1], s.v[fc.v2]);
        if (area <= 0.0f) continue;
        float le = luminance(mt.emissive.x * mt.emissive_strength, mt.emissive.y * mt.emissive_strength,
                             mt.emissive.z * mt.emissive_strength);
        float wt = area * std::fmax(le, 0.0f);
        if (wt <= 0.0f) continue;
        s.emissive.push_back({i, 0.0f, 0.0f});
        w.push_back(wt);
    }
    if (!s.emissive.empty()) {
        float sum = 0.0f;
        for (float x : w) sum += x;
        if (sum <= 0.0f) {
            s.emissive.clear();
        } else {
            float acc = 0.0f;
            for (size_t i = 0; i < s.emissive.size(); ++i) {
                float pmf = w[i] / sum;
                acc += pmf;
                s.emissive[i].pmf = pmf;
                s.emissive[i].cdf = acc;
                s.triPmf[s.emissive[i].tri] = pmf;
            }
            s.emissive.back().cdf = 1.0f;
            s.triAlias = buildAliasTable(w);
        }
    }
    // :112-149
    s.envPmf.clear();
    s.envAlias.clear();
    if (s.env.w > 0 && s.env.h > 0) {
        const int n = s.env.w * s.env.h;
        std::vector<float> ew(static_cast<size_t>(n), 0.0f);
        for (int y = 0; y < s.env.h; ++y) {
            float theta = 3.14159265358979323846f * (static_cast<float>(y) + 0.5f) / static_cast<float>(s.env.h);
            float st = std::fmax(std::sin(theta), 0.0f);
            for (int x = 0; x < s.env.w; ++x) {
                size_t idx = static_cast<size_t>(y) * s.env.w + x;
                const float* c = &s.env.rgba[idx * 4];
                ew[idx] = std::fmax(luminance(c[0], c[1], c[2]), 0.0f) * st;
            }
        }
        float sum = 0.0f;
        for (float x : ew) sum += x;
        if (sum > 0.0f) {
            s.envPmf.resize(static_cast<size_t>(n));
            for (int i = 0; i < n; ++i) s.envPmf[i] = ew[i] / sum;
            s.envAlias = buildAliasTable(ew);
        }
    }
}

// ------------------------------------------------ texture fetch (CUDA semantics)
// cuda/cuda_texture.cu:39-63 creates float4 textures with linear filtering,
// normalized coordinates and GL-derived address modes.  Linear filtering as
// documented for the texture unit: xB = u*W - 0.5, i = floor(xB), weights
// from frac(xB) kept with 8 fractional bits.
inline int foldIndex(long long i, int n, int mode) {
    if (mode == 1) return static_cast<int>(i < 0 ? 0 : (i >= n ? n - 1 : i));
    if (mode == 2) {
        long long p = 2LL * n, m = ((i % p) + p) % p;
        return static_cast<int>(m < n ? m : p - 1 - m);
    }
    return static_cast<int>(((i % n) + n) % n);
}
inline int wrapMode(int gl) { return gl == 33071 ? 1 : (gl == 33648 ? 2 : 0); }

// The filter weights are those of the texture unit (float coordinates, 8-bit
// fractions); the weighted sum is formed in T (float: the renderer's
// arithmetic; double: the FP64 oracle's).
template <class T>
void tex2DT(const float* rgba, int w, int h, int ws, int wt, float u, float v, T out[4]) {
    if (ws == 0) u = u - std::floor(u);
    if (wt == 0) v = v - std::floor(v);
    float x = u * static_cast<float>(w) - 0.5f, y = v * static_cast<float>(h) - 0.5f;
    float fx = std::floor(x), fy = std::floor(y);
    float a = std::nearbyint((x - fx) * 256.0f) / 256.0f, b = std::nearbyint((y - fy) * 256.0f) / 256.0f;
    if (!std::isfinite(a)) a = 0.0f;
    if (!std::isfinite(b)) b = 0.0f;
    long long x0 = std::isfinite(fx) ? static_cast<long long>(std::fmax(std::fmin(fx, 1e15f), -1e15f)) : 0;
    long long y0 = std::isfinite(fy) ? static_cast<long long>(std::fmax(std::fmin(fy, 1e15f), -1e15f)) : 0;
    int xa = foldIndex(x0, w, ws), xb = foldIndex(x0 + 1, w, ws);
    int ya = foldIndex(y0, h, wt), yb = foldIndex(y0 + 1, h, wt);
    const float* t00 = rgba + 4 * (static_cast<size_t>(ya) * w + xa);
    const float* t10 = rgba + 4 * (static_cast<size_t>(ya) * w + xb);
    const float* t01 = rgba + 4 * (static_cast<size_t>(yb) * w + xa);
    const float* t11 = rgba + 4 * (static_cast<size_t>(yb) * w + xb);
    const float w00 = (1 - a) * (1 - b), w10 = a * (1 - b), w01 = (1 - a) * b, w11 = a * b;
    for (int c = 0; c < 4; ++c)
        out[c] = T(w00) * paramRead<T>(t00 + c) + T(w10) * paramRead<T>(t10 + c) + T(w01) * paramRead<T>(t01 + c) +
                 T(w11) * paramRead<T>(t11 + c);
}
void tex2D(const float* rgba, int w, int h, int ws, int wt, float u, float v, float out[4]) {
    tex2DT<float>(rgba, w, h, ws, wt, u, v, out);
}

// -------------------------------------------------------- materials/microfacet.cuh
template <class R> R D_GGX(R noh, R alpha) {  // :10-14
    R a2 = alpha * alpha;
    R d = noh * noh * (a2 - R(1)) + R(1);
    return a2 / (K<R>::Pi * d * d + R(1e-7));
}
template <class R> R D_GGX_Aniso(R noh, R hx, R hy, R ax, R ay) {  // :17-22
    R x = hx / ax, y = hy / ay;
    R d = x * x + y * y + noh * noh;
    return R(1) / (K<R>::Pi * ax * ay * d * d + R(1e-7));
}
template <class R> R G1_Smith(R nov, R alpha) {  // :25-28
    R a2 = alpha * alpha;
    return R(2) * nov / (nov + msqrt(a2 + (R(1) - a2) * nov * nov) + R(1e-7));
}
template <class R> R G2_Smith(R nol, R nov, R alpha) {  // :31-36
    R a2 = alpha * alpha;
    R gv = nol * msqrt(a2 + (R(1) - a2) * nov * nov);
    R gl = nov * msqrt(a2 + (R(1) - a2) * nol * nol);
    return R(2) * nol * nov / (gv + gl + R(1e-7));
}
template <class R> R F_SchlickS(R c, R f0) {  // :39-43
    R t = clampr(R(1) - c, R(0), R(1));
    R t2 = t * t;
    return f0 + (R(1) - f0) * t2 * t2 * t;
}
template <class R> V3<R> F_SchlickV(R c, const V3<R>& f0) {  // :46-51
    R t = clampr(R(1) - c, R(0), R(1));
    R t2 = t * t;
    R t5 = t2 * t2 * t;
    return f0 + (V3<R>(1) - f0) * t5;
}
template <class R> R F_Dielectric(R ci0, R eta) {  // :54-62
    R s2t = eta * eta * (R(1) - ci0 * ci0);
    if (s2t >= R(1)) return R(1);
    R ct = msqrt(fmx(R(0), R(1) - s2t));
    R ci = mfabs(ci0);
    R rs = (ci - eta * ct) / (ci + eta * ct + R(1e-7));
    R rp = (eta * ci - ct) / (eta * ci + ct + R(1e-7));
    return R(0.5) * (rs * rs + rp * rp);
}
template <class R> V3<R> sampleVNDF(const V3<R>& wo, R ax, R ay, R u1, R u2) {  // :67-92
    V3<R> wh = normalize(V3<R>(ax * wo.x, ay * wo.y, wo.z));
    if (wh.z < 0) wh = -wh;
    V3<R> T1 = (wh.z < R(0.9999)) ? normalize(cross(V3<R>(0, 0, 1), wh)) : V3<R>(1, 0, 0);
    V3<R> T2 = cross(wh, T1);
    R r = msqrt(u1), phi = K<R>::TwoPi * u2;
    R t1 = r * std::cos(phi), t2 = r * std::sin(phi);
    R s = R(0.5) * (R(1) + wh.z);
    t2 = (R(1) - s) * msqrt(fmx(R(0), R(1) - t1 * t1)) + s * t2;
    V3<R> nh = t1 * T1 + t2 * T2 + msqrt(fmx(R(0), R(1) - t1 * t1 - t2 * t2)) * wh;
    return normalize(V3<R>(ax * nh.x, ay * nh.y, fmx(R(0), nh.z)));
}
template <class R> R pdfVNDF(const V3<R>& wo, const V3<R>& h, R ax, R ay) {  // :95-102
    R nov = fmx(wo.z, R(1e-5)), noh = fmx(h.z, R(0)), voh = fmx(dot(wo, h), R(0));
    R d = D_GGX_Aniso(noh, h.x, h.y, ax, ay);
    R g1 = G1_Smith(nov, msqrt(ax * ay));
    return d * g1 * voh / (nov + R(1e-7));
}
template <class R> bool refractDir(const V3<R>& wi, const V3<R>& n, R eta, V3<R>& wt) {  // :105-113
    R ci = dot(wi, n);
    R s2i = fmx(R(0), R(1) - ci * ci);
    R s2t = eta * eta * s2i;
    if (s2t >= R(1)) return false;
    R ct = msqrt(R(1) - s2t);
    wt = eta * (-wi) + (eta * ci - ct) * n;
    return true;
}

// ------------------------------------------------------ materials/disney_bsdf.cuh
template <class R>
struct MatEval {  // :20-36
    V3<R> base;
    R metallic, roughness, specTrans, ior, clearcoat, clearcoatRoughness, sheen, sheenTint, subsurface;
    R anisotropic, specularTint, specularFactor;
    V3<R> specularColor, sheenColor, emissive, attColor;
    R emissiveStrength, anisotropyRotation, thickness, attDistance;
};

template <class R> V3<R> powColor(const R c[4], bool linear) {
    if (linear) return V3<R>(c[0], c[1], c[2]);
    return V3<R>(mpow(c[0], R(2.2)), mpow(c[1], R(2.2)), mpow(c[2], R(2.2)));
}

template <class R>
MatEval<R> fetchMaterial(const oracle_scene* s, const pt_material& m, float u, float v) {  // :53-122
    MatEval<R> e;
    e.base = V3<R>(paramRead<R>(&m.base_color.x), paramRead<R>(&m.base_color.y), paramRead<R>(&m.base_color.z));
    e.metallic = paramRead<R>(&m.metallic); e.roughness = paramRead<R>(&m.roughness); e.specTrans = R(m.spec_trans); e.ior = R(m.ior);
    e.clearcoat = R(m.clearcoat); e.clearcoatRoughness = R(m.clearcoat_roughness);
    e.sheen = R(m.sheen); e.sheenTint = R(m.sheen_tint); e.subsurface = R(m.subsurface);
    e.anisotropic = R(m.anisotropic); e.specularTint = R(m.specular_tint); e.specularFactor = R(m.specular_factor);
    e.specularColor = V3<R>(R(m.specular_color_r), R(m.specular_color_g), R(m.specular_color_b));
    e.emissive = v3of<R>(m.emissive);
    e.emissiveStrength = R(m.emissive_strength);
    e.sheenColor = V3<R>(R(m.sheen_color_r), R(m.sheen_color_g), R(m.sheen_color_b));
    e.anisotropyRotation = R(m.anisotropy_rotation);
    e.thickness = R(m.thickness_factor);
    e.attDistance = R(m.attenuation_distance);
    e.attColor = V3<R>(R(m.att_color_r), R(m.att_color_g), R(m.att_color_b));
    if (s && !s->tex.empty()) {
        auto valid = [&](int id) { return id >= 0 && static_cast<size_t>(id) < s->tex.size(); };
        auto fetch = [&](int id, R o[4]) {
            const Texture& t = s->tex[id];
            tex2DT<R>(t.rgba.data(), t.w, t.h, t.wrapS, t.wrapT, u, v, o);
        };
        R c[4];
        if (valid(m.base_color_tex_id)) { fetch(m.base_color_tex_id, c); e.base = e.base * powColor<R>(c, s->tex[m.base_color_tex_id].linear); }
        if (valid(m.metallic_roughness_tex_id)) { fetch(m.metallic_roughness_tex_id, c); e.roughness *= c[1]; e.metallic *= c[2]; }
        if (valid(m.emissive_tex_id)) { fetch(m.emissive_tex_id, c); e.emissive = e.emissive * powColor<R>(c, s->tex[m.emissive_tex_id].linear); }
        if (valid(m.clearcoat_tex_id)) { fetch(m.clearcoat_tex_id, c); e.clearcoat *= R(c[0]); }
        if (valid(m.clearcoat_roughness_tex_id)) { fetch(m.clearcoat_roughness_tex_id, c); e.clearcoatRoughness *= R(c[1]); }
        if (valid(m.transmission_tex_id)) { fetch(m.transmission_tex_id, c); e.specTrans *= R(c[0]); }
        if (valid(m.sheen_color_tex_id)) { fetch(m.sheen_color_tex_id, c); e.sheenColor = e.sheenColor * powColor<R>(c, s->tex[m.sheen_color_tex_id].linear); }
        if (valid(m.sheen_roughness_tex_id)) { fetch(m.sheen_roughness_tex_id, c); e.sheen *= R(c[3]); }
    }
    e.roughness = clampr(e.roughness, R(0.001), R(1));
    e.metallic = clampr(e.metallic, R(0), R(1));
    e.clearcoatRoughness = clampr(e.clearcoatRoughness, R(0.001), R(1));
    return e;
}

// interpolateTangent + applyNormalMap (materials/shading_utils.cuh:29-68):
// the tangent-space normal of the normal texture (scaled in x/y by
// normalScale) in the frame of the interpolated mesh tangent with the first
// vertex's handedness, or in Frame(N) when the mesh has no tangent there.
template <class R>
V3<R> normalMap(const oracle_scene* s, const pt_material& m, const V3<R>& N, const pt_vertex& va,
                const pt_vertex& vb, const pt_vertex& vc, R bu, R bv, float u, float v) {
    if (m.normal_tex_id < 0 || static_cast<size_t>(m.normal_tex_id) >= s->tex.size()) return N;
    const Texture& t = s->tex[m.normal_tex_id];
    R nm[4];
    tex2DT<R>(t.rgba.data(), t.w, t.h, t.wrapS, t.wrapT, u, v, nm);
    V3<R> mapN(nm[0] * R(2) - R(1), nm[1] * R(2) - R(1), nm[2] * R(2) - R(1));
    mapN.x *= R(m.normal_scale);
    mapN.y *= R(m.normal_scale);
    mapN = normalize(mapN);
    const R w = R(1) - bu - bv;
    V3<R> T = normalize(V3<R>(w * R(va.tangent[0]) + bu * R(vb.tangent[0]) + bv * R(vc.tangent[0]),
                              w * R(va.tangent[1]) + bu * R(vb.tangent[1]) + bv * R(vc.tangent[1]),
                              w * R(va.tangent[2]) + bu * R(vb.tangent[2]) + bv * R(vc.tangent[2])));
    const R sign = va.tangent[3] >= 0.0f ? R(1) : R(-1);
    if (T.x != R(0) || T.y != R(0) || T.z != R(0)) {
        T = normalize(T - N * dot(N, T));
        V3<R> B = cross(N, T) * sign;
        return normalize(T * mapN.x + B * mapN.y + N * mapN.z);
    }
    Frame<R> tbn(N);
    return normalize(tbn.toWorld(mapN));
}

template <class R> void computeAlpha(R rough, R aniso, R& ax, R& ay) {  // :126-131
    R aspect = msqrt(R(1) - R(0.9) * aniso);
    R a = rough * rough;
    ax = fmx(a / aspect, R(0.001));
    ay = fmx(a * aspect, R(0.001));
}
template <class R> R lum(const V3<R>& c) { return R(0.2126) * c.x + R(0.7152) * c.y + R(0.0722) * c.z; }
template <class R> V3<R> evalDisneyDiffuse(const MatEval<R>& m, R nol, R nov, R hol) {  // :135-159
    R fd90 = R(0.5) + R(2) * hol * hol * m.roughness;
    R fl = R(1) + (fd90 - R(1)) * mpow(R(1) - nol, R(5));
    R fv = R(1) + (fd90 - R(1)) * mpow(R(1) - nov, R(5));
    V3<R> d = m.base * (K<R>::InvPi * fl * fv);
    if (m.subsurface > R(0)) {
        R f90 = hol * hol * m.roughness;
        R sl = R(1) + (f90 - R(1)) * mpow(R(1) - nol, R(5));
        R sv = R(1) + (f90 - R(1)) * mpow(R(1) - nov, R(5));
        R ss = R(1.25) * (sl * sv * (R(1) / (nol + nov + R(1e-5)) - R(0.5)) + R(0.5));
        V3<R> sd = m.base * (K<R>::InvPi * ss);
        d = (R(1) - m.subsurface) * d + m.subsurface * sd;
    }
    return d;
}
template <class R> V3<R> evalSheen(const MatEval<R>& m, R hol) {  // :163-171
    if (m.sheen <= R(0)) return V3<R>(0);
    R l = lum(m.base);
    V3<R> tint = l > R(0) ? m.base / l : V3<R>(1);
    V3<R> bs = (R(1) - m.sheenTint) * V3<R>(1) + m.sheenTint * tint;
    V3<R> fc = bs * m.sheenColor;
    R fh = mpow(R(1) - hol, R(5));
    return m.sheen * fc * fh;
}
template <class R> V3<R> computeF0(const MatEval<R>& m) {  // :175-183
    R l = lum(m.base);
    V3<R> tint = l > R(0) ? m.base / l : V3<R>(1);
    R eta = m.ior;
    R r0 = ((eta - R(1)) * (eta - R(1))) / ((eta + R(1)) * (eta + R(1)));
    V3<R> spec = r0 * ((R(1) - m.specularTint) * V3<R>(1) + m.specularTint * tint);
    spec = spec * m.specularFactor * m.specularColor;
    return (R(1) - m.metallic) * spec + m.metallic * m.base;
}
template <class R> void lobeWeights(const MatEval<R>& m, R& wd, R& ws, R& wc, R& wt, R& wsh) {  // :187-205
    R dw = (R(1) - m.metallic) * (R(1) - m.specTrans);
    R sw = R(1);
    R tw = (R(1) - m.metallic) * m.specTrans;
    R cw = R(0.25) * m.clearcoat;
    R shw = (R(1) - m.metallic) * m.sheen * R(0.1);
    R total = dw + sw + cw + tw + shw;
    if (total <= R(0)) { wd = ws = wc = wt = wsh = R(0); return; }
    R inv = R(1) / total;
    wd = dw * inv; ws = sw * inv; wc = cw * inv; wt = tw * inv; wsh = shw * inv;
}

template <class R> struct BsdfEval { V3<R> f; R pdf = R(0); };

template <class R>
BsdfEval<R> evalDisneyBsdf(const MatEval<R>& m, const V3<R>& wo, const V3<R>& wi, const V3<R>& N, bool entering) {
    // :209-321
    BsdfEval<R> r;
    R nol = dot(N, wi), nov = dot(N, wo);
    bool refl = nol > R(0);
    if (!refl && m.specTrans <= R(0)) return r;
    if (!refl) {
        R anl = mfabs(nol), anv = mfabs(nov);
        if (anl < R(1e-5) || anv < R(1e-5)) return r;
        R eta = entering ? (R(1) / m.ior) : m.ior;
        V3<R> ht = -normalize(wo * eta + wi);
        R hov = mfabs(dot(ht, wo));
        R ax, ay;
        computeAlpha(m.roughness, m.anisotropic, ax, ay);
        Frame<R> fr(N);
        V3<R> hl = fr.toLocal(ht);
        R D = D_GGX_Aniso(mfabs(hl.z), hl.x, hl.y, ax, ay);
        R G = G2_Smith(anl, anv, msqrt(ax * ay));
        R F = F_Dielectric(hov, eta);
        R sd = dot(wo, ht) * eta + dot(wi, ht);
        R factor = mfabs(dot(wo, ht)) * mfabs(dot(wi, ht)) / (anv * anl * sd * sd + R(1e-7));
        R tw = (R(1) - m.metallic) * m.specTrans;
        V3<R> ft = m.base * tw * D * G * (R(1) - F) * factor;
        R wd, ws, wc, wt, wsh;
        lobeWeights(m, wd, ws, wc, wt, wsh);
        R jac = eta * eta * mfabs(dot(wi, ht)) / (sd * sd + R(1e-7));
        R tp = D * G1_Smith(anv, msqrt(ax * ay)) * hov / (anv + R(1e-7)) * jac;
        r.f = ft;
        r.pdf = wt * tp;
        return r;
    }
    if (nol < R(1e-5) || nov < R(1e-5)) return r;
    V3<R> H = normalize(wo + wi);
    R noh = fmx(dot(N, H), R(0)), hol = fmx(dot(H, wi), R(0));
    Frame<R> fr(N);
    V3<R> hl = fr.toLocal(H);
    R ax, ay;
    computeAlpha(m.roughness, m.anisotropic, ax, ay);
    R wd, ws, wc, wt, wsh;
    lobeWeights(m, wd, ws, wc, wt, wsh);
    V3<R> f(0);
    R pdf = R(0);
    R dw = (R(1) - m.metallic) * (R(1) - m.specTrans);
    if (dw > R(0)) {
        f += dw * evalDisneyDiffuse(m, nol, nov, hol);
        pdf += wd * (nol * K<R>::InvPi);
    }
    if (m.sheen > R(0) && m.metallic < R(1)) {
        f += (R(1) - m.metallic) * evalSheen(m, hol);
        pdf += wsh * (nol * K<R>::InvPi);
    }
    {
        R D = D_GGX_Aniso(noh, hl.x, hl.y, ax, ay);
        R G = G2_Smith(nol, nov, msqrt(ax * ay));
        V3<R> F = F_SchlickV(hol, computeF0(m));
        f += D * G * F / (R(4) * nol * nov + R(1e-7));
        pdf += ws * (pdfVNDF(fr.toLocal(wo), hl, ax, ay) / (R(4) * hol + R(1e-7)));
    }
    if (m.clearcoat > R(0)) {
        R ca = fmx(m.clearcoatRoughness * m.clearcoatRoughness, R(0.001));
        R Dc = D_GGX(noh, ca);
        R Gc = G2_Smith(nol, nov, ca);
        R Fc = F_SchlickS(hol, R(0.04));
        R cs = R(0.25) * m.clearcoat * Dc * Gc * Fc / (R(4) * nol * nov + R(1e-7));
        f += V3<R>(cs);
        pdf += wc * (Dc * noh / (R(4) * hol + R(1e-7)));
    }
    r.f = f * nol;
    r.pdf = pdf;
    return r;
}

template <class R> struct BsdfSample { V3<R> wi, weight; R pdf = R(0); bool delta = false, valid = false; };

// :325-410; the two numbers of each warp are drawn left to right (u1 first).
template <class R, class Rng>
BsdfSample<R> sampleDisneyBsdf(const MatEval<R>& m, const V3<R>& wo, const V3<R>& N, Rng& rng, bool entering) {
    BsdfSample<R> res;
    R nov = dot(N, wo);
    if (nov < R(1e-5)) return res;
    R wd, ws, wc, wt, wsh;
    lobeWeights(m, wd, ws, wc, wt, wsh);
    R choice = R(rng.get1D());
    Frame<R> fr(N);
    V3<R> wol = fr.toLocal(wo);
    R ax, ay;
    computeAlpha(m.roughness, m.anisotropic, ax, ay);
    V3<R> wi;
    bool isDelta = false;
    if (choice < wd + wsh) {
        R u1 = R(rng.get1D());
        R u2 = R(rng.get1D());
        wi = fr.toWorld(sampleCosineHemisphere(u1, u2));
    } else if (choice < wd + wsh + ws) {
        R u1 = R(rng.get1D());
        R u2 = R(rng.get1D());
        V3<R> H = fr.toWorld(sampleVNDF(wol, ax, ay, u1, u2));
        wi = reflect(-wo, H);
        if (dot(wi, N) <= R(0)) return res;
        isDelta = m.roughness < R(0.01);
    } else if (choice < wd + wsh + ws + wc) {
        R ca = fmx(m.clearcoatRoughness * m.clearcoatRoughness, R(0.001));
        R u1 = R(rng.get1D());
        R u2 = R(rng.get1D());
        V3<R> H = fr.toWorld(sampleVNDF(wol, ca, ca, u1, u2));
        wi = reflect(-wo, H);
        if (dot(wi, N) <= R(0)) return res;
    } else {
        R u1 = R(rng.get1D());
        R u2 = R(rng.get1D());
        V3<R> H = fr.toWorld(sampleVNDF(wol, ax, ay, u1, u2));
        R eta = entering ? (R(1) / m.ior) : m.ior;
        V3<R> wt2;
        if (!refractDir(wo, H, eta, wt2)) {
            wi = reflect(-wo, H);
            if (dot(wi, N) <= R(0)) return res;
        } else {
            wi = normalize(wt2);
            isDelta = m.roughness < R(0.01);
        }
    }
    BsdfEval<R> ev = evalDisneyBsdf(m, wo, wi, N, entering);
    if (ev.pdf <= R(0)) return res;
    res.wi = wi;
    res.pdf = ev.pdf;
    res.delta = isDelta;
    res.valid = true;
    if (isDelta) {
        if (dot(N, wi) < R(0)) res.weight = m.base;
        else {
            V3<R> H = normalize(wo + wi);
            res.weight = F_SchlickV(fmx(dot(H, wi), R(0)), computeF0(m));
        }
    } else {
        res.weight = ev.f / (ev.pdf + R(1e-7));
    }
    return res;
}

// -------------------------------------------- geometry/triangle.h + traversal
template <class R>
struct Hit {
    bool hit = false;
    R t = R(FLT_MAX);
    R u = 0, v = 0;
    int prim = -1;
};

template <class R>
R triHit(const V3<R>& o, const V3<R>& d, const pt_vertex& a, const pt_vertex& b, const pt_vertex& c, R& u, R& v) {
    // triangle.h:13-51 (returns t or -1; the t >= 1e-4 test is the caller's)
    V3<R> p0 = v3of<R>(a.position);
    V3<R> e1 = v3of<R>(b.position) - p0, e2 = v3of<R>(c.position) - p0;
    V3<R> pv = cross(d, e2);
    R det = dot(e1, pv);
    if (mfabs(det) < R(FLT_EPSILON)) return R(-1);
    R inv = R(1) / det;
    V3<R> tv = o - p0;
    u = dot(tv, pv) * inv;
    if (u < R(0) || u > R(1)) return R(-1);
    V3<R> qv = cross(tv, e1);
    v = dot(d, qv) * inv;
    if (v < R(0) || u + v > R(1)) return R(-1);
    return dot(e2, qv) * inv;
}

template <class R>
bool box(const V3<R>& o, const V3<R>& inv, const Node& n, R tmax) {  // bvh_traversal.cuh:29-41
    R tmin = R(0);
    for (int i = 0; i < 3; ++i) {
        R t1 = (R(n.bmin[i]) - o[i]) * inv[i], t2 = (R(n.bmax[i]) - o[i]) * inv[i];
        tmin = fmn(fmx(t1, tmin), fmx(t2, tmin));
        tmax = fmx(fmn(t1, tmax), fmn(t2, tmax));
    }
    return tmin <= tmax;
}

template <class R>
Hit<R> intersectScene(const oracle_scene* s, const V3<R>& o, const V3<R>& d, uint64_t* nodeCount = nullptr) {
    // bvh_traversal.cuh:43-96 (the stack grows as needed instead of 32 fixed entries)
    Hit<R> res;
    if (s->nodes.empty()) return res;
    V3<R> inv(R(1) / d.x, R(1) / d.y, R(1) / d.z);
    int sign[3] = {inv.x < 0 ? 1 : 0, inv.y < 0 ? 1 : 0, inv.z < 0 ? 1 : 0};
    std::vector<uint32_t> st;
    st.reserve(64);
    st.push_back(0);
    while (!st.empty()) {
        uint32_t idx = st.back();
        st.pop_back();
        const Node& n = s->nodes[idx];
        if (nodeCount) ++*nodeCount;
        if (!box(o, inv, n, res.t)) continue;
        if (n.isLeaf()) {
            for (uint32_t i = 0; i < n.data1; ++i) {
                uint32_t p = n.first() + i;
                const pt_face& f = s->f[p];
                R u, v;
                R t = triHit(o, d, s->v[f.v0], s->v[f.v1], s->v[f.v2], u, v);
                if (t < R(1e-4)) continue;  // also rejects misses (-1)
                if (t < res.t) { res.hit = true; res.t = t; res.u = u; res.v = v; res.prim = static_cast<int>(p); }
            }
        } else {
            uint32_t l = n.data0, r = n.data1;
            int axis = n.maxDim();
            float cl = s->nodes[l].centroid(axis), cr = s->nodes[r].centroid(axis);
            bool flip = sign[axis] != 0;
            if (flip == (cl < cr)) { st.push_back(l); st.push_back(r); }
            else { st.push_back(r); st.push_back(l); }
        }
    }
    return res;
}

template <class R>
bool occludedScene(const oracle_scene* s, const V3<R>& o, const V3<R>& d, R tmax) {  // :98-133
    if (s->nodes.empty()) return false;
    V3<R> inv(R(1) / d.x, R(1) / d.y, R(1) / d.z);
    std::vector<uint32_t> st;
    st.reserve(64);
    st.push_back(0);
    while (!st.empty()) {
        uint32_t idx = st.back();
        st.pop_back();
        const Node& n = s->nodes[idx];
        if (!box(o, inv, n, tmax)) continue;
        if (n.isLeaf()) {
            for (uint32_t i = 0; i < n.data1; ++i) {
                const pt_face& f = s->f[n.first() + i];
                R u, v;
                R t = triHit(o, d, s->v[f.v0], s->v[f.v1], s->v[f.v2], u, v);
                if (t > R(1e-4) && t < tmax) return true;
            }
        } else {
            st.push_back(n.data0);
            st.push_back(n.data1);
        }
    }
    return false;
}

// ------------------------------------------------------------ lights/envmap.h
template <class R> void dirToUv(const V3<R>& d, R& u, R& v) {  // :15-24
    u = std::atan2(d.z, d.x);
    v = std::asin(d.y);
    u *= K<R>::Inv2Pi;
    v *= K<R>::InvPi;
    u += R(0.5);
    v += R(0.5);
    v = R(1) - v;
}
template <class R> V3<R> uvToDir(R u, R v) {  // :27-32
    R phi = (u - R(0.5)) * K<R>::TwoPi, theta = v * K<R>::Pi;
    R st = std::sin(theta);
    return normalize(V3<R>(std::cos(phi) * st, std::cos(theta), std::sin(phi) * st));
}
template <class R> V3<R> envRadiance(const oracle_scene* s, const V3<R>& d) {  // :34-45 (x 2 pi)
    if (s->env.w <= 0) return V3<R>(0);
    R u, v;
    dirToUv(d, u, v);
    R c[4];
    tex2DT<R>(s->env.rgba.data(), s->env.w, s->env.h, s->env.wrapS, s->env.wrapT, static_cast<float>(u),
              static_cast<float>(v), c);
    return V3<R>(c[0], c[1], c[2]) * K<R>::TwoPi;
}
template <class R> R envPdf(const oracle_scene* s, const V3<R>& d) {  // light_pdf.cuh:12-27
    if (s->env.w <= 0 || s->envPmf.empty()) return R(0);
    R u, v;
    dirToUv(d, u, v);
    int x = static_cast<int>(clampr(u * R(s->env.w), R(0), R(s->env.w - 1)));
    int y = static_cast<int>(clampr(v * R(s->env.h), R(0), R(s->env.h - 1)));
    R st = msqrt(fmx(R(0), R(1) - d.y * d.y));
    if (st <= R(1e-5)) return R(0);
    return R(s->envPmf[static_cast<size_t>(y) * s->env.w + x]) * R(s->env.w * s->env.h) / (K<R>::TwoPi * K<R>::Pi * st);
}
template <class R> R emissivePdf(const oracle_scene* s, int prim, R dist, const V3<R>& d) {  // light_pdf.cuh:29-54
    if (prim < 0 || prim >= static_cast<int>(s->f.size())) return R(0);
    R pmf = R(s->triPmf[prim]);
    if (pmf <= R(0)) return R(0);
    const pt_face& f = s->f[prim];
    V3<R> p0 = v3of<R>(s->v[f.v0].position), p1 = v3of<R>(s->v[f.v1].position), p2 = v3of<R>(s->v[f.v2].position);
    V3<R> c = cross(p1 - p0, p2 - p0);
    R area = R(0.5) * length(c);
    if (area <= R(0)) return R(0);
    R cl = mfabs(dot(normalize(c), -d));
    if (cl <= R(1e-5)) return R(0);
    return pmf * dist * dist / (area * cl);
}

// --------------------------------------------- integrators/direct_lighting.cuh
template <class R> struct LightSample { V3<R> wi, L; R pdf = R(0); bool valid = false; };

struct Ctx {
    uint64_t closest = 0, shadow = 0;
};

template <class R>
bool isOccluded(const oracle_scene* s, const V3<R>& p, const V3<R>& n, const V3<R>& dir, R maxd, Ctx& c) {  // :36-51
    R sign = dot(dir, n) >= R(0) ? R(1) : R(-1);
    V3<R> o = p + n * (R(1e-4) * sign);
    R tmax = (maxd < R(FLT_MAX)) ? (maxd - R(1e-3)) : R(FLT_MAX);
    c.shadow++;
    return occludedScene(s, o, normalize(dir), tmax);
}

inline uint32_t sampleAlias(const std::vector<AliasEntry>& t, uint32_t n, float u1, float u2) {  // :55-59
    uint32_t idx = static_cast<uint32_t>(u1 * n);
    if (idx >= n) idx = n - 1;
    return (u2 < t[idx].prob) ? idx : t[idx].alias;
}

template <class R>
LightSample<R> sampleAreaLight(const oracle_scene* s, const V3<R>& p, const V3<R>& n, Sampler& rng, Ctx& c) {
    // :75-121
    LightSample<R> ls;
    if (s->emissive.empty()) return ls;
    float a1 = rng.get1D();
    float a2 = rng.get1D();
    const auto& ref = s->emissive[sampleAlias(s->triAlias, static_cast<uint32_t>(s->emissive.size()), a1, a2)];
    const pt_face& f = s->f[ref.tri];
    const pt_vertex &va = s->v[f.v0], &vb = s->v[f.v1], &vc = s->v[f.v2];
    R t1 = R(rng.get1D());
    R t2 = R(rng.get1D());
    R su0 = msqrt(t1);  // light_sampler.h:58-71
    R b0 = R(1) - su0, b1 = t2 * su0, b2 = R(1) - b0 - b1;
    V3<R> lp = v3of<R>(va.position) * b0 + v3of<R>(vb.position) * b1 + v3of<R>(vc.position) * b2;
    V3<R> ln = normalize(v3of<R>(va.normal) * b0 + v3of<R>(vb.normal) * b1 + v3of<R>(vc.normal) * b2);
    V3<R> e1 = v3of<R>(vb.position) - v3of<R>(va.position), e2 = v3of<R>(vc.position) - v3of<R>(va.position);
    if (length(ln) == R(0)) ln = normalize(cross(e1, e2));
    V3<R> tl = lp - p;
    R d2 = dot(tl, tl);
    if (d2 <= R(1e-8)) return ls;
    R dist = msqrt(d2);
    V3<R> wi = tl / dist;
    if (fmx(R(0), dot(n, wi)) <= R(0)) return ls;
    R area = R(0.5) * length(cross(e1, e2));
    if (area <= R(0)) return ls;
    R cl = mfabs(dot(ln, -wi));
    if (cl <= R(1e-5)) return ls;
    R pdf = (R(ref.pmf) / area) * d2 / cl;
    if (pdf <= R(0)) return ls;
    if (isOccluded(s, p, n, wi, dist, c)) return ls;
    const pt_material& lm = s->m[f.material_id];
    ls.wi = wi;
    ls.L = v3of<R>(lm.emissive) * R(lm.emissive_strength);
    ls.pdf = pdf;
    ls.valid = true;
    return ls;
}

template <class R>
LightSample<R> sampleEnvLight(const oracle_scene* s, const V3<R>& p, const V3<R>& n, Sampler& rng, Ctx& c) {
    // :123-154
    LightSample<R> ls;
    float a1 = rng.get1D();
    float a2 = rng.get1D();
    uint32_t texel = sampleAlias(s->envAlias, static_cast<uint32_t>(s->envPmf.size()), a1, a2);
    int x = static_cast<int>(texel % static_cast<uint32_t>(s->env.w));
    int y = static_cast<int>(texel / static_cast<uint32_t>(s->env.w));
    R j1 = R(rng.get1D());
    R j2 = R(rng.get1D());
    R u = (R(x) + j1) / R(s->env.w), v = (R(y) + j2) / R(s->env.h);
    V3<R> wi = uvToDir(u, v);
    if (fmx(R(0), dot(n, wi)) <= R(0)) return ls;
    R pdf = envPdf(s, wi);
    if (pdf <= R(0)) return ls;
    if (isOccluded(s, p, n, wi, R(FLT_MAX), c)) return ls;
    ls.wi = wi;
    ls.L = envRadiance(s, wi);
    ls.pdf = pdf;
    ls.valid = true;
    return ls;
}

template <class R>
V3<R> evalDirect(const LightSample<R>& ls, const MatEval<R>& m, const V3<R>& wo, const V3<R>& n) {  // :204-218
    if (!ls.valid) return V3<R>(0);
    BsdfEval<R> b = evalDisneyBsdf(m, wo, ls.wi, n, true);
    if (b.pdf <= R(0)) return V3<R>(0);
    R mis = powerHeuristic(ls.pdf, b.pdf);
    return ls.L * b.f * (mis / (ls.pdf + R(1e-7)));
}

template <class R>
V3<R> punctualLights(const oracle_scene* s, const V3<R>& p, const MatEval<R>& m, const V3<R>& wo, const V3<R>& n,
                     Ctx& c) {  // :158-200
    V3<R> ld(0);
    for (const auto& l : s->lights) {
        V3<R> wi;
        R dist = R(FLT_MAX);
        V3<R> I;
        if (l.type == 0 || l.type == 2) {
            V3<R> tl = v3of<R>(l.position) - p;
            R d2 = dot(tl, tl);
            if (d2 < R(1e-8)) continue;
            dist = msqrt(d2);
            wi = tl / dist;
            R att = dist >= R(l.range) ? R(0) : R(1) / (dist * dist + R(1e-7));  // punctual_light.h:36-39
            if (l.type == 2) {                                                  // :28-34
                R ca = dot(normalize(-tl), v3of<R>(l.direction));
                R sp;
                if (ca < R(l.outer_cone)) sp = R(0);
                else if (ca > R(l.inner_cone)) sp = R(1);
                else {
                    R t = (ca - R(l.outer_cone)) / (R(l.inner_cone) - R(l.outer_cone) + R(1e-7));
                    sp = t * t;
                }
                att *= sp;
            }
            I = v3of<R>(l.color) * (R(l.intensity) * att);
        } else {
            wi = -v3of<R>(l.direction);
            I = v3of<R>(l.color) * R(l.intensity);
        }
        if (dot(n, wi) <= R(0)) continue;
        if (isOccluded(s, p, n, wi, dist, c)) continue;
        ld += I * evalDisneyBsdf(m, wo, wi, n, true).f;
    }
    return ld;
}

// ------------------------------------------------------- integrators/path.cu
// The shading point of a closest hit (path.cu:86-103): position, the
// interpolated normal and uv (triangle.h:40-49), the entering test with the
// double-sided flip, the uv transform (shading_utils.cuh:17-27) and the
// normal map.  The reference's tangent flip for double-sided back faces
// (path.cu:101-102) tests `!entering && doubleSided` after the flip above
// has set entering, so it never fires and the tangent is used as
// interpolated.
template <class R>
struct SurfacePt {
    V3<R> P, sn, gn, N, wo;
    R tu, tv;      // interpolated uv (the alpha test reads them untransformed)
    float mu, mv;  // after the uv transform
    bool entering;
    const pt_material* mat;
};
template <class R>
SurfacePt<R> surfaceAt(const oracle_scene* s, const Hit<R>& h, const V3<R>& o, const V3<R>& d) {
    SurfacePt<R> sp;
    const pt_face& f = s->f[h.prim];
    const pt_vertex &va = s->v[f.v0], &vb = s->v[f.v1], &vc = s->v[f.v2];
    R w = R(1) - h.u - h.v;
    sp.P = o + h.t * d;
    sp.sn = normalize(v3of<R>(va.normal) * w + v3of<R>(vb.normal) * h.u + v3of<R>(vc.normal) * h.v);
    sp.tu = R(va.u) * w + R(vb.u) * h.u + R(vc.u) * h.v;
    sp.tv = R(va.v) * w + R(vb.v) * h.u + R(vc.v) * h.v;
    const pt_material& mat = s->m[f.material_id];
    sp.mat = &mat;
    sp.entering = dot(sp.sn, d) < R(0);
    if (mat.double_sided && !sp.entering) { sp.gn = -sp.sn; sp.entering = true; }
    else sp.gn = sp.entering ? sp.sn : -sp.sn;
    sp.wo = -d;
    float mu = static_cast<float>(sp.tu), mv = static_cast<float>(sp.tv);
    if (!(mat.uv_rotation == 0.0f && mat.uv_scale_x == 1.0f && mat.uv_scale_y == 1.0f && mat.uv_offset_x == 0.0f &&
          mat.uv_offset_y == 0.0f)) {  // shading_utils.cuh:17-27
        float cr = std::cos(mat.uv_rotation), sr = std::sin(mat.uv_rotation);
        float ru = cr * mu + sr * mv, rv = -sr * mu + cr * mv;
        mu = ru * mat.uv_scale_x + mat.uv_offset_x;
        mv = rv * mat.uv_scale_y + mat.uv_offset_y;
    }
    sp.mu = mu;
    sp.mv = mv;
    sp.N = normalMap<R>(s, mat, sp.gn, va, vb, vc, h.u, h.v, mu, mv);
    return sp;
}

template <class R>
V3<R> pathTrace(const oracle_scene* s, const pt_camera_frame& cf, int W, int H, int maxDepth, int x, int y,
                Sampler& rng, Ctx& c) {
    V3<R> Lo(0), T(1);
    // emitRay :31-59
    R jx = R(rng.get1D());
    R jy = R(rng.get1D());
    R u = (R(2) * (R(x) + jx) / R(W) - R(1)) * R(cf.tan_half_fov) * R(cf.aspect);
    R v = (R(2) * (R(y) + jy) / R(H) - R(1)) * R(cf.tan_half_fov);
    V3<R> pos(R(cf.position[0]), R(cf.position[1]), R(cf.position[2]));
    V3<R> fwd(R(cf.forward[0]), R(cf.forward[1]), R(cf.forward[2]));
    V3<R> rgt(R(cf.right[0]), R(cf.right[1]), R(cf.right[2]));
    V3<R> cup(R(cf.up[0]), R(cf.up[1]), R(cf.up[2]));
    V3<R> dir = normalize(u * rgt + v * cup + fwd);
    if (cf.aperture > 0.0f) {
        V3<R> focal = pos + dir * (R(cf.focal_distance) / dot(dir, fwd));
        R l1 = R(rng.get1D());
        R l2 = R(rng.get1D());
        R r = msqrt(l1) * R(cf.aperture);
        R th = K<R>::TwoPi * l2;
        pos = pos + (rgt * (r * std::cos(th)) + cup * (r * std::sin(th)));
        dir = normalize(focal - pos);
    }
    V3<R> o = pos, d = normalize(dir);
    R lastPdf = 0;
    bool lastDelta = true;
    const bool hasArea = !s->emissive.empty();
    const bool hasEnv = s->env.w > 0 && !s->envPmf.empty();
    for (int depth = 0; depth < maxDepth; ++depth) {  // :69-199
        c.closest++;
        Hit<R> h = intersectScene(s, o, d);
        if (!h.hit) {
            V3<R> env = envRadiance(s, d);
            if (env != V3<R>(0)) {
                R mis = R(1);
                if (depth > 0 && !lastDelta) mis = powerHeuristic(lastPdf, envPdf(s, d));
                Lo += T * env * mis;
            }
            break;
        }
        const SurfacePt<R> sp = surfaceAt(s, h, o, d);
        const pt_material& mat = *sp.mat;
        const V3<R>&P = sp.P, &gn = sp.gn, &wo = sp.wo, &N = sp.N;
        const R tu = sp.tu, tv = sp.tv;
        const bool entering = sp.entering;
        MatEval<R> me = fetchMaterial<R>(s, mat, sp.mu, sp.mv);
        if (isEmissive(mat)) {
            R mis = R(1);
            if (depth > 0 && !lastDelta) mis = powerHeuristic(lastPdf, emissivePdf(s, h.prim, h.t, d));
            Lo += T * (v3of<R>(mat.emissive) * R(mat.emissive_strength)) * mis;
            break;
        }
        if (mat.alpha_mode == 1 || mat.alpha_mode == 2) {  // :135-158
            float alpha = mat.alpha_mode == 1 ? 1.0f : mat.base_alpha;
            int tid = mat.base_color_tex_id;
            if (tid >= 0 && static_cast<size_t>(tid) < s->tex.size()) {
                float cc[4];
                const Texture& t = s->tex[tid];
                tex2D(t.rgba.data(), t.w, t.h, t.wrapS, t.wrapT, static_cast<float>(tu), static_cast<float>(tv), cc);
                alpha = mat.alpha_mode == 1 ? cc[3] : alpha * cc[3];
            }
            bool skip = mat.alpha_mode == 1 ? (alpha < mat.alpha_cutoff) : (rng.get1D() >= alpha);
            if (skip) {
                o = P + d * R(1e-4);
                d = normalize(d);
                continue;
            }
        }
        bool deltaSurf = (me.metallic >= R(1) && me.roughness < R(0.01)) || (me.specTrans >= R(1) && me.roughness < R(0.01));
        if (!deltaSurf) {  // :161-166
            V3<R> ld(0);
            if (hasArea && hasEnv) {
                ld += evalDirect(sampleAreaLight<R>(s, P, N, rng, c), me, wo, N);
                ld += evalDirect(sampleEnvLight<R>(s, P, N, rng, c), me, wo, N);
            } else if (hasArea) {
                ld = evalDirect(sampleAreaLight<R>(s, P, N, rng, c), me, wo, N);
            } else if (hasEnv) {
                ld = evalDirect(sampleEnvLight<R>(s, P, N, rng, c), me, wo, N);
            }
            Lo += T * ld;
            Lo += T * punctualLights(s, P, me, wo, N, c);
        }
        BsdfSample<R> bs = sampleDisneyBsdf(me, wo, N, rng, entering);
        if (!bs.valid) break;
        T *= bs.weight;
        bool trans = dot(bs.wi, gn) < R(0);
        V3<R> off = trans ? -gn : gn;
        o = P + off * R(1e-4);
        d = normalize(bs.wi);
        if (trans && me.thickness > R(0) && me.attDistance < R(1e20)) {  // :178-187
            R t = me.thickness, ad = me.attDistance;
            V3<R> ac = me.attColor;
            T *= V3<R>(std::exp(-t * fmx(-std::log(fmx(ac.x, R(1e-6))), R(0)) / ad),
                       std::exp(-t * fmx(-std::log(fmx(ac.y, R(1e-6))), R(0)) / ad),
                       std::exp(-t * fmx(-std::log(fmx(ac.z, R(1e-6))), R(0)) / ad));
        }
        lastPdf = bs.pdf;
        lastDelta = bs.delta;
        if (depth > 2) {  // :193-198
            R surv = fmn(fmx(fmx(T.x, T.y), T.z), R(0.95));
            if (R(rng.get1D()) >= surv) break;
            T /= surv;
        }
    }
    Lo.x = fmn(Lo.x, R(100));  // :201-204
    Lo.y = fmn(Lo.y, R(100));
    Lo.z = fmn(Lo.z, R(100));
    return Lo;
}

// ----------------------------------------------- detached estimator (adjoint)
// A path traced at the current parameters is recorded as a tape of its
// decisions (hit points, sampled and light directions, pdfs, MIS weights,
// Russian-roulette factors).  Evaluating the tape with other material /
// texture / environment values gives the "detached" estimator L(theta):
// directions and pdfs stay those of the recorded path, while BSDF values,
// base colours and environment radiance follow theta.  Its derivative is
// what a path-replay adjoint estimates, so central differences of it check
// an adjoint path by path.
struct TapeVertex {
    int kind = 2;  // 0 environment on a miss, 1 emitter hit, 2 surface, 3 alpha pass-through
    V3<float> d;   // kind 0: direction
    double mis = 1.0;
    V3<double> Le;  // kind 1: emitted radiance
    int mat = 0;
    float mu = 0, mv = 0;
    V3<float> wo, N;
    bool entering = true;
    bool nee = false, hasA = false, hasE = false;
    V3<float> wiA, wiE;
    V3<double> LA;
    double kA = 0, kE = 0;  // mis / (pdf_light + 1e-7) at the recorded parameters
    std::vector<std::pair<V3<float>, V3<double>>> punct;  // visible punctual lights: direction, intensity
    bool valid = false, delta = false;  // BSDF continuation
    V3<float> wi;
    double pdf0 = 0;
    V3<double> beer = V3<double>(1.0);
    double surv = 1.0;
    bool rrDead = false;
};

inline V3<double> d3(const V3<float>& v) { return V3<double>(v.x, v.y, v.z); }

// pathTrace<float> with the tape recorded (same draws, same decisions).
std::vector<TapeVertex> pathTape(const oracle_scene* s, const pt_camera_frame& cf, int W, int H, int maxDepth, int x,
                                 int y, Sampler& rng, Ctx& c) {
    using R = float;
    std::vector<TapeVertex> tape;
    V3<R> T(1);
    R jx = R(rng.get1D());
    R jy = R(rng.get1D());
    R u = (R(2) * (R(x) + jx) / R(W) - R(1)) * R(cf.tan_half_fov) * R(cf.aspect);
    R v = (R(2) * (R(y) + jy) / R(H) - R(1)) * R(cf.tan_half_fov);
    V3<R> pos(cf.position[0], cf.position[1], cf.position[2]);
    V3<R> fwd(cf.forward[0], cf.forward[1], cf.forward[2]);
    V3<R> rgt(cf.right[0], cf.right[1], cf.right[2]);
    V3<R> cup(cf.up[0], cf.up[1], cf.up[2]);
    V3<R> dir = normalize(u * rgt + v * cup + fwd);
    if (cf.aperture > 0.0f) {
        V3<R> focal = pos + dir * (R(cf.focal_distance) / dot(dir, fwd));
        R l1 = R(rng.get1D());
        R l2 = R(rng.get1D());
        R r = msqrt(l1) * R(cf.aperture);
        R th = K<R>::TwoPi * l2;
        pos = pos + (rgt * (r * std::cos(th)) + cup * (r * std::sin(th)));
        dir = normalize(focal - pos);
    }
    V3<R> o = pos, d = normalize(dir);
    R lastPdf = 0;
    bool lastDelta = true;
    const bool hasArea = !s->emissive.empty();
    const bool hasEnv = s->env.w > 0 && !s->envPmf.empty();
    for (int depth = 0; depth < maxDepth; ++depth) {
        c.closest++;
        Hit<R> h = intersectScene(s, o, d);
        TapeVertex tv;
        if (!h.hit) {
            V3<R> env = envRadiance(s, d);
            if (env != V3<R>(0)) {
                R mis = R(1);
                if (depth > 0 && !lastDelta) mis = powerHeuristic(lastPdf, envPdf(s, d));
                tv.kind = 0;
                tv.d = d;
                tv.mis = mis;
                tape.push_back(tv);
            }
            break;
        }
        const SurfacePt<R> sp = surfaceAt(s, h, o, d);
        const pt_material& mat = *sp.mat;
        const V3<R>&P = sp.P, &gn = sp.gn, &wo = sp.wo, &N = sp.N;
        const R tu = sp.tu, tvv = sp.tv;
        const bool entering = sp.entering;
        const float mu = sp.mu, mv = sp.mv;
        MatEval<R> me = fetchMaterial<R>(s, mat, mu, mv);
        if (isEmissive(mat)) {
            R mis = R(1);
            if (depth > 0 && !lastDelta) mis = powerHeuristic(lastPdf, emissivePdf(s, h.prim, h.t, d));
            tv.kind = 1;
            tv.Le = d3(v3of<R>(mat.emissive) * R(mat.emissive_strength));
            tv.mis = mis;
            tape.push_back(tv);
            break;
        }
        if (mat.alpha_mode == 1 || mat.alpha_mode == 2) {
            float alpha = mat.alpha_mode == 1 ? 1.0f : mat.base_alpha;
            int tid = mat.base_color_tex_id;
            if (tid >= 0 && static_cast<size_t>(tid) < s->tex.size()) {
                float cc[4];
                const Texture& t = s->tex[tid];
                tex2D(t.rgba.data(), t.w, t.h, t.wrapS, t.wrapT, tu, tvv, cc);
                alpha = mat.alpha_mode == 1 ? cc[3] : alpha * cc[3];
            }
            bool skip = mat.alpha_mode == 1 ? (alpha < mat.alpha_cutoff) : (rng.get1D() >= alpha);
            if (skip) {
                tv.kind = 3;
                tape.push_back(tv);
                o = P + d * R(1e-4);
                d = normalize(d);
                continue;
            }
        }
        tv.kind = 2;
        tv.mat = static_cast<int>(sp.mat - s->m.data());
        tv.mu = mu;
        tv.mv = mv;
        tv.wo = wo;
        tv.N = N;
        tv.entering = entering;
        bool deltaSurf = (me.metallic >= R(1) && me.roughness < R(0.01)) || (me.specTrans >= R(1) && me.roughness < R(0.01));
        if (!deltaSurf) {
            tv.nee = true;
            auto record = [&](const LightSample<R>& ls, bool env) {
                if (!ls.valid) return;
                BsdfEval<R> b = evalDisneyBsdf(me, wo, ls.wi, N, true);
                if (b.pdf <= R(0)) return;
                double k = double(powerHeuristic(ls.pdf, b.pdf)) / (double(ls.pdf) + 1e-7);
                if (env) { tv.hasE = true; tv.wiE = ls.wi; tv.kE = k; }
                else { tv.hasA = true; tv.wiA = ls.wi; tv.LA = d3(ls.L); tv.kA = k; }
            };
            if (hasArea) record(sampleAreaLight<R>(s, P, N, rng, c), false);
            if (hasEnv) record(sampleEnvLight<R>(s, P, N, rng, c), true);
            for (const auto& l : s->lights) {  // evaluatePunctualLights' visibility, recorded
                V3<R> wi;
                R dist = R(FLT_MAX);
                V3<R> I;
                if (l.type == 0 || l.type == 2) {
                    V3<R> tl = v3of<R>(l.position) - P;
                    R d2 = dot(tl, tl);
                    if (d2 < R(1e-8)) continue;
                    dist = msqrt(d2);
                    wi = tl / dist;
                    R att = dist >= R(l.range) ? R(0) : R(1) / (dist * dist + R(1e-7));
                    if (l.type == 2) {
                        R ca = dot(normalize(-tl), v3of<R>(l.direction));
                        R sp;
                        if (ca < R(l.outer_cone)) sp = R(0);
                        else if (ca > R(l.inner_cone)) sp = R(1);
                        else {
                            R t = (ca - R(l.outer_cone)) / (R(l.inner_cone) - R(l.outer_cone) + R(1e-7));
                            sp = t * t;
                        }
                        att *= sp;
                    }
                    I = v3of<R>(l.color) * (R(l.intensity) * att);
                } else {
                    wi = -v3of<R>(l.direction);
                    I = v3of<R>(l.color) * R(l.intensity);
                }
                if (dot(N, wi) <= R(0)) continue;
                if (isOccluded(s, P, N, wi, dist, c)) continue;
                tv.punct.push_back({wi, d3(I)});
            }
        }
        BsdfSample<R> bs = sampleDisneyBsdf(me, wo, N, rng, entering);
        if (!bs.valid) {
            tape.push_back(tv);
            break;
        }
        tv.valid = true;
        tv.delta = bs.delta;
        tv.wi = bs.wi;
        tv.pdf0 = bs.pdf;
        T *= bs.weight;
        bool trans = dot(bs.wi, gn) < R(0);
        V3<R> off = trans ? -gn : gn;
        o = P + off * R(1e-4);
        d = normalize(bs.wi);
        if (trans && me.thickness > R(0) && me.attDistance < R(1e20)) {
            R t = me.thickness, ad = me.attDistance;
            V3<R> ac = me.attColor;
            V3<R> bf(std::exp(-t * fmx(-std::log(fmx(ac.x, R(1e-6))), R(0)) / ad),
                     std::exp(-t * fmx(-std::log(fmx(ac.y, R(1e-6))), R(0)) / ad),
                     std::exp(-t * fmx(-std::log(fmx(ac.z, R(1e-6))), R(0)) / ad));
            T *= bf;
            tv.beer = d3(bf);
        }
        lastPdf = bs.pdf;
        lastDelta = bs.delta;
        if (depth > 2) {
            R surv = fmn(fmx(fmx(T.x, T.y), T.z), R(0.95));
            tv.surv = surv;
            if (R(rng.get1D()) >= surv) {
                tv.rrDead = true;
                tape.push_back(tv);
                break;
            }
            T /= surv;
        }
        tape.push_back(tv);
    }
    return tape;
}

// ------------------------------------------ reverse-mode type (host adjoint)
// AD is a double with an index into a per-thread list of nodes (-1: a
// constant).  Every operation on a non-constant records its operands and
// local partial derivatives; reads of registered parameters (material base
// colour, roughness, metallic; texels of textures and of the environment)
// become leaves.  One backward sweep over the list then gives the derivative
// of a path's radiance with respect to every parameter the path read.  The
// integrator's templates run on AD unchanged (their math calls go through
// msqrt / mpow / mfabs / fmn / fmx, overloaded here); a min or max passes the
// derivative of the operand it returns, the first one on a tie.
struct ADNode {
    int32_t a, b;
    double da, db;
};
struct ADTape {
    std::vector<ADNode> nodes;
    std::vector<std::pair<int32_t, int64_t>> leaves;  // node, parameter id
    // parameter arrays: materials (5 ids each), then texel channels of every
    // texture in order, then of the environment
    const pt_material* mats = nullptr;
    size_t nmat = 0;
    struct Block { const float* p; size_t n; int64_t base; };
    std::vector<Block> blocks;
    int64_t nparams = 0;
    int64_t paramId(const float* p) const {
        const char* c = reinterpret_cast<const char*>(p);
        const char* m0 = reinterpret_cast<const char*>(mats);
        if (mats && c >= m0 && c < m0 + nmat * sizeof(pt_material)) {
            const size_t off = static_cast<size_t>(c - m0), i = off / sizeof(pt_material);
            const size_t f = off % sizeof(pt_material);
            int k = -1;
            if (f == offsetof(pt_material, base_color) + 0) k = 0;
            else if (f == offsetof(pt_material, base_color) + 4) k = 1;
            else if (f == offsetof(pt_material, base_color) + 8) k = 2;
            else if (f == offsetof(pt_material, roughness)) k = 3;
            else if (f == offsetof(pt_material, metallic)) k = 4;
            return k < 0 ? -1 : static_cast<int64_t>(5 * i + k);
        }
        for (const Block& b : blocks)
            if (p >= b.p && p < b.p + b.n) return b.base + (p - b.p);
        return -1;
    }
};
thread_local ADTape* tlTape = nullptr;

struct AD {
    double v;
    int32_t i;
    constexpr AD() : v(0.0), i(-1) {}
    constexpr AD(double x) : v(x), i(-1) {}
    AD(double x, int32_t idx) : v(x), i(idx) {}
};
inline AD adOp(double v, int32_t a, double da, int32_t b, double db) {
    if (a < 0 && b < 0) return AD(v);
    tlTape->nodes.push_back({a, b, da, db});
    return AD(v, static_cast<int32_t>(tlTape->nodes.size() - 1));
}
inline AD operator+(AD a, AD b) { return adOp(a.v + b.v, a.i, 1.0, b.i, 1.0); }
inline AD operator-(AD a, AD b) { return adOp(a.v - b.v, a.i, 1.0, b.i, -1.0); }
inline AD operator*(AD a, AD b) { return adOp(a.v * b.v, a.i, b.v, b.i, a.v); }
inline AD operator/(AD a, AD b) {
    const double q = a.v / b.v;
    return adOp(q, a.i, 1.0 / b.v, b.i, -q / b.v);
}
inline AD operator-(AD a) { return adOp(-a.v, a.i, -1.0, -1, 0.0); }
inline AD& operator+=(AD& a, AD b) { return a = a + b; }
inline AD& operator-=(AD& a, AD b) { return a = a - b; }
inline AD& operator*=(AD& a, AD b) { return a = a * b; }
inline AD& operator/=(AD& a, AD b) { return a = a / b; }
inline bool operator<(AD a, AD b) { return a.v < b.v; }
inline bool operator>(AD a, AD b) { return a.v > b.v; }
inline bool operator<=(AD a, AD b) { return a.v <= b.v; }
inline bool operator>=(AD a, AD b) { return a.v >= b.v; }
inline bool operator==(AD a, AD b) { return a.v == b.v; }
inline bool operator!=(AD a, AD b) { return a.v != b.v; }
inline AD msqrt(AD a) {
    const double r = std::sqrt(a.v);
    return adOp(r, a.i, r > 0.0 ? 0.5 / r : 0.0, -1, 0.0);
}
inline AD mpow(AD a, AD b) {
    const double r = std::pow(a.v, b.v);
    const double da = a.v != 0.0 ? b.v * std::pow(a.v, b.v - 1.0) : 0.0;
    const double db = a.v > 0.0 ? r * std::log(a.v) : 0.0;
    return adOp(r, a.i, da, b.i, db);
}
inline AD mfabs(AD a) { return a.v < 0.0 ? -a : a; }
inline AD fmn(AD a, AD b) { return (b.v < a.v || std::isnan(a.v)) ? b : a; }
inline AD fmx(AD a, AD b) { return (b.v > a.v || std::isnan(a.v)) ? b : a; }
template <>
inline AD paramRead<AD>(const float* p) {
    const int64_t id = tlTape ? tlTape->paramId(p) : -1;
    if (id < 0) return AD(*p);
    tlTape->nodes.push_back({-1, -1, 0.0, 0.0});
    const int32_t n = static_cast<int32_t>(tlTape->nodes.size() - 1);
    tlTape->leaves.push_back({n, id});
    return AD(*p, n);
}

template <class R> V3<R> lift(const V3<double>& a) { return V3<R>(R(a.x), R(a.y), R(a.z)); }

// sampleEnvironmentRadiance (envmap.h:34-45, x 2 pi) at a fixed direction,
// the texels in R.
template <class R> V3<R> envRadianceAt(const oracle_scene* s, const V3<double>& d) {
    if (s->env.w <= 0) return V3<R>(0);
    double u, v;
    dirToUv(d, u, v);
    R c[4];
    tex2DT<R>(s->env.rgba.data(), s->env.w, s->env.h, s->env.wrapS, s->env.wrapT, static_cast<float>(u),
              static_cast<float>(v), c);
    return V3<R>(c[0], c[1], c[2]) * K<R>::TwoPi;
}

// The detached estimator of a tape under the parameters of scene sp (no
// clamp): in double for the finite differences, in AD for the adjoint.
template <class R>
V3<R> tapeEvalT(const oracle_scene* sp, const std::vector<TapeVertex>& tape) {
    V3<R> L(0), T(1);
    for (const TapeVertex& v : tape) {
        if (v.kind == 0) {
            L += T * envRadianceAt<R>(sp, d3(v.d)) * R(v.mis);
            break;
        }
        if (v.kind == 1) {
            L += T * lift<R>(v.Le) * R(v.mis);
            break;
        }
        if (v.kind == 3) continue;
        MatEval<R> me = fetchMaterial<R>(sp, sp->m[v.mat], v.mu, v.mv);
        const V3<R> wo = lift<R>(d3(v.wo)), N = lift<R>(d3(v.N));
        if (v.nee) {
            V3<R> ld(0);
            if (v.hasA) ld += lift<R>(v.LA) * evalDisneyBsdf(me, wo, lift<R>(d3(v.wiA)), N, true).f * R(v.kA);
            if (v.hasE)
                ld += envRadianceAt<R>(sp, d3(v.wiE)) * evalDisneyBsdf(me, wo, lift<R>(d3(v.wiE)), N, true).f * R(v.kE);
            L += T * ld;
            V3<R> pl(0);
            for (const auto& p : v.punct) pl += lift<R>(p.second) * evalDisneyBsdf(me, wo, lift<R>(d3(p.first)), N, true).f;
            L += T * pl;
        }
        if (!v.valid) break;
        const V3<R> wi = lift<R>(d3(v.wi));
        V3<R> weight;
        if (v.delta) {
            if (dot(N, wi) < R(0.0)) weight = me.base;
            else weight = F_SchlickV(fmx(dot(normalize(wo + wi), wi), R(0.0)), computeF0(me));
        } else {
            weight = evalDisneyBsdf(me, wo, wi, N, v.entering).f / R(v.pdf0 + 1e-7);
        }
        T *= weight;
        T *= lift<R>(v.beer);
        if (v.rrDead) break;
        T /= R(v.surv);
    }
    return L;
}
inline V3<double> tapeEval(const oracle_scene* sp, const std::vector<TapeVertex>& tape) {
    return tapeEvalT<double>(sp, tape);
}

}  // namespace orc

// =============================================================== C interface
using namespace orc;

extern "C" {

oracle_scene* oracle_scene_create(const pt_vertex* v, uint32_t nv, const pt_face* f, uint32_t nf, const pt_material* m,
                                  uint32_t nm, const float* env, int ew, int eh, const pt_punctual_light* lights,
                                  uint32_t nl, int bins, int leaf) {
    oracle_scene* s = new oracle_scene();
    s->v.assign(v, v + nv);
    s->f.assign(f, f + nf);
    s->m.assign(m, m + nm);
    if (lights && nl) s->lights.assign(lights, lights + nl);
    if (env && ew > 0 && eh > 0) {
        s->env.w = ew;
        s->env.h = eh;
        s->env.rgba.assign(env, env + static_cast<size_t>(ew) * eh * 4);
    }
    Builder b;
    b.buckets = bins > 0 ? bins : 24;
    b.leafThreshold = leaf > 0 ? leaf : 4;
    b.build(s->v, s->f, s->order);
    s->nodes = b.nodes;
    buildLightTables(*s);
    return s;
}

int oracle_scene_add_texture(oracle_scene* s, const float* rgba, int w, int h, int ws, int wt, int linear) {
    Texture t;
    t.w = w; t.h = h; t.wrapS = wrapMode(ws); t.wrapT = wrapMode(wt); t.linear = linear;
    t.rgba.assign(rgba, rgba + static_cast<size_t>(w) * h * 4);
    s->tex.push_back(std::move(t));
    return static_cast<int>(s->tex.size() - 1);
}

void oracle_scene_destroy(oracle_scene* s) { delete s; }

uint32_t oracle_bvh_num_nodes(const oracle_scene* s) { return static_cast<uint32_t>(s->nodes.size()); }

void oracle_bvh_get(const oracle_scene* s, float* nodes8, uint32_t* order) {
    if (nodes8) std::memcpy(nodes8, s->nodes.data(), s->nodes.size() * sizeof(Node));
    if (order) std::memcpy(order, s->order.data(), s->order.size() * 4);
}

void oracle_trace(const oracle_scene* s, const float* rays8, int n, int any, float* t, int32_t* face, float* bary) {
    for (int i = 0; i < n; ++i) {
        const float* r = rays8 + 8 * static_cast<size_t>(i);
        V3<float> o(r[0], r[1], r[2]);
        V3<float> d = normalize(V3<float>(r[3], r[4], r[5]));  // core/ray.h:13
        if (any) {
            t[i] = occludedScene(s, o, d, r[6]) ? 1.0f : 0.0f;
            face[i] = -1;
            continue;
        }
        Hit<float> h = intersectScene(s, o, d);
        t[i] = h.t;
        face[i] = h.hit ? static_cast<int32_t>(s->order[h.prim]) : -1;
        if (bary) { bary[2 * i] = h.u; bary[2 * i + 1] = h.v; }
    }
}

void oracle_light_tables(const oracle_scene* s, uint32_t* ne, uint32_t* nenv, float* tp, float* ta, float* ep,
                         float* ea) {
    if (ne) *ne = static_cast<uint32_t>(s->emissive.size());
    if (nenv) *nenv = static_cast<uint32_t>(s->envPmf.size());
    if (tp) std::memcpy(tp, s->triPmf.data(), s->triPmf.size() * 4);
    if (ta) std::memcpy(ta, s->triAlias.data(), s->triAlias.size() * 8);
    if (ep) std::memcpy(ep, s->envPmf.data(), s->envPmf.size() * 4);
    if (ea) std::memcpy(ea, s->envAlias.data(), s->envAlias.size() * 8);
}

uint32_t oracle_alias_table(const float* w, uint32_t n, float* prob, uint32_t* alias) {
    std::vector<float> ww(w, w + n);
    auto t = buildAliasTable(ww);
    for (size_t i = 0; i < t.size(); ++i) { prob[i] = t[i].prob; alias[i] = t[i].alias; }
    return static_cast<uint32_t>(t.size());
}

uint32_t oracle_sobol_u32(int px, int py, uint32_t index, uint32_t dim, uint32_t seed) {
    return sobolU32(px, py, index, dim, seed);
}
float oracle_sobol_float(int px, int py, uint32_t index, uint32_t dim, uint32_t seed) {
    return sobolFloat(sobolU32(px, py, index, dim, seed));
}
void oracle_pcg(int x, int y, uint32_t frame, uint32_t sample, uint32_t seed, int n, uint32_t* u32, float* f) {
    Sampler a, b;
    a.initPCG(x, y, frame, sample, seed);
    b.initPCG(x, y, frame, sample, seed);
    for (int i = 0; i < n; ++i) {
        if (u32) u32[i] = a.nextU32();
        if (f) f[i] = b.get1D();
    }
}

void oracle_camera_frame(const pt_camera* c, pt_camera_frame* out) {  // camera/camera.h:36-51
    V3<float> fwd = normalize(v3of<float>(c->look_at) - v3of<float>(c->position));
    V3<float> rgt = normalize(cross(fwd, v3of<float>(c->up)));
    V3<float> cup = cross(rgt, fwd);
    std::memset(out, 0, sizeof(*out));
    out->position[0] = c->position.x; out->position[1] = c->position.y; out->position[2] = c->position.z;
    out->forward[0] = fwd.x; out->forward[1] = fwd.y; out->forward[2] = fwd.z;
    out->right[0] = rgt.x; out->right[1] = rgt.y; out->right[2] = rgt.z;
    out->up[0] = cup.x; out->up[1] = cup.y; out->up[2] = cup.z;
    out->tan_half_fov = std::tan(c->fov * 3.14159265358979323846f / 360.0f);
    out->aspect = static_cast<float>(c->width) / c->height;
    out->aperture = c->aperture;
    out->focal_distance = c->focal_distance;
}

void oracle_bsdf_eval(const pt_material* m, const double wo[3], const double wi[3], const double n[3], int entering,
                      int fp64, double f[3], double* pdf) {
    if (fp64) {
        MatEval<double> me = fetchMaterial<double>(nullptr, *m, 0, 0);
        BsdfEval<double> e = evalDisneyBsdf(me, V3<double>(wo[0], wo[1], wo[2]), V3<double>(wi[0], wi[1], wi[2]),
                                            V3<double>(n[0], n[1], n[2]), entering != 0);
        f[0] = e.f.x; f[1] = e.f.y; f[2] = e.f.z; *pdf = e.pdf;
    } else {
        MatEval<float> me = fetchMaterial<float>(nullptr, *m, 0, 0);
        BsdfEval<float> e = evalDisneyBsdf(me, V3<float>(float(wo[0]), float(wo[1]), float(wo[2])),
                                           V3<float>(float(wi[0]), float(wi[1]), float(wi[2])),
                                           V3<float>(float(n[0]), float(n[1]), float(n[2])), entering != 0);
        f[0] = e.f.x; f[1] = e.f.y; f[2] = e.f.z; *pdf = e.pdf;
    }
}

namespace {
struct FixedRng {
    const float* u;
    int i = 0;
    float get1D() { return u[i++]; }
};
}  // namespace

int oracle_bsdf_sample(const pt_material* m, const double wo[3], const double n[3], int entering, const float u[3],
                       double wi[3], double w[3], double* pdf, int* delta) {
    MatEval<float> me = fetchMaterial<float>(nullptr, *m, 0, 0);
    FixedRng rng{u};
    BsdfSample<float> b = sampleDisneyBsdf(me, V3<float>(float(wo[0]), float(wo[1]), float(wo[2])),
                                           V3<float>(float(n[0]), float(n[1]), float(n[2])), rng, entering != 0);
    wi[0] = b.wi.x; wi[1] = b.wi.y; wi[2] = b.wi.z;
    w[0] = b.weight.x; w[1] = b.weight.y; w[2] = b.weight.z;
    *pdf = b.pdf;
    *delta = b.delta ? 1 : 0;
    return b.valid ? 1 : 0;
}

void oracle_dir_to_uv(const float d[3], float uv[2]) { dirToUv(V3<float>(d[0], d[1], d[2]), uv[0], uv[1]); }
void oracle_uv_to_dir(const float uv[2], float d[3]) {
    V3<float> r = uvToDir(uv[0], uv[1]);
    d[0] = r.x; d[1] = r.y; d[2] = r.z;
}
float oracle_env_pdf(const oracle_scene* s, const float d[3]) { return envPdf(s, V3<float>(d[0], d[1], d[2])); }

double oracle_emissive_pdf(const oracle_scene* s, int prim, double dist, const double d[3]) {
    return emissivePdf<double>(s, prim, dist, V3<double>(d[0], d[1], d[2]));
}

int oracle_sample_light(const oracle_scene* s, int env, const double p[3], const double n[3], int x, int y,
                        uint32_t frame, uint32_t sample, double wi[3], double L[3], double* pdf) {
    Sampler rng;
    rng.initPCG(x, y, frame, sample, 0u);
    Ctx c;
    const V3<double> P(p[0], p[1], p[2]), N(n[0], n[1], n[2]);
    LightSample<double> ls = env ? sampleEnvLight<double>(s, P, N, rng, c) : sampleAreaLight<double>(s, P, N, rng, c);
    if (!ls.valid) return 0;
    wi[0] = ls.wi.x, wi[1] = ls.wi.y, wi[2] = ls.wi.z;
    L[0] = ls.L.x, L[1] = ls.L.y, L[2] = ls.L.z;
    *pdf = ls.pdf;
    return 1;
}
void oracle_fetch_material(const oracle_scene* s, const pt_material* m, float u, float v, float out[20]) {
    const MatEval<float> e = fetchMaterial<float>(s, *m, u, v);
    const float vals[20] = {e.base.x, e.base.y, e.base.z, e.metallic, e.roughness, e.specTrans, e.ior, e.clearcoat,
                            e.clearcoatRoughness, e.sheen, e.sheenTint, e.subsurface, e.anisotropic,
                            e.sheenColor.x, e.sheenColor.y, e.sheenColor.z, e.emissive.x, e.emissive.y, e.emissive.z,
                            e.specularFactor};
    std::memcpy(out, vals, sizeof(vals));
}

int oracle_surface(const oracle_scene* s, const float o3[3], const float d3v[3], float out[16]) {
    const V3<float> o(o3[0], o3[1], o3[2]);
    const V3<float> d = normalize(V3<float>(d3v[0], d3v[1], d3v[2]));  // core/ray.h:13
    const Hit<float> h = intersectScene(s, o, d);
    std::memset(out, 0, 16 * sizeof(float));
    if (!h.hit) return 0;
    const SurfacePt<float> sp = surfaceAt(s, h, o, d);
    const float vals[16] = {1.0f, h.t, sp.P.x, sp.P.y, sp.P.z, sp.sn.x, sp.sn.y, sp.sn.z, sp.gn.x, sp.gn.y, sp.gn.z,
                            sp.N.x, sp.N.y, sp.N.z, sp.entering ? 1.0f : 0.0f,
                            static_cast<float>(s->order[h.prim])};
    std::memcpy(out, vals, sizeof(vals));
    return 1;
}

void oracle_tex2d(const float* rgba, int w, int h, int ws, int wt, float u, float v, float out[4]) {
    tex2D(rgba, w, h, wrapMode(ws), wrapMode(wt), u, v, out);
}

void oracle_render(const oracle_scene* s, const pt_camera* cam, int spf, int maxDepth, int sampler, uint32_t seed,
                   uint32_t frame, int x0, int x1, int y0, int y1, int threads, int fp64, float* out,
                   uint64_t* counters) {
    pt_camera_frame cf;
    oracle_camera_frame(cam, &cf);
    const int W = cam->width, H = cam->height;
    const int rw = x1 - x0, rh = y1 - y0;
    if (rw <= 0 || rh <= 0) return;
    int nt = threads > 0 ? threads : static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
    std::atomic<int> next{0};
    std::atomic<uint64_t> cl{0}, sh{0}, paths{0};
    auto work = [&]() {
        Ctx c;
        uint64_t np = 0;
        while (true) {
            int row = next.fetch_add(1);
            if (row >= rh) break;
            int y = y0 + row;
            for (int x = x0; x < x1; ++x) {
                double ad[3] = {0, 0, 0};
                float af[3] = {0, 0, 0};
                for (int i = 0; i < spf; ++i) {  // path.cu:216-228
                    Sampler rng;
                    if (sampler == 1) rng.initSobol(x, y, frame * static_cast<uint32_t>(spf) + static_cast<uint32_t>(i), seed);
                    else rng.initPCG(x, y, frame, static_cast<uint32_t>(i), seed);
                    if (fp64) {
                        V3<double> L = pathTrace<double>(s, cf, W, H, maxDepth, x, y, rng, c);
                        ad[0] += L.x; ad[1] += L.y; ad[2] += L.z;
                    } else {
                        V3<float> L = pathTrace<float>(s, cf, W, H, maxDepth, x, y, rng, c);
                        af[0] += L.x; af[1] += L.y; af[2] += L.z;
                    }
                    ++np;
                }
                float* o = out + 3 * (static_cast<size_t>(row) * rw + (x - x0));
                for (int k = 0; k < 3; ++k)
                    o[k] = fp64 ? static_cast<float>(ad[k] / spf) : af[k] / static_cast<float>(spf);
            }
        }
        cl += c.closest;
        sh += c.shadow;
        paths += np;
    };
    std::vector<std::thread> pool;
    for (int i = 1; i < nt; ++i) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
    if (counters) {
        counters[0] = cl.load();
        counters[1] = sh.load();
        counters[2] = paths.load();
    }
}

void oracle_microfacet(const float* in, const float* f0, const float* wo3, const float* n3, float* out) {
    using R = float;
    const R noh = in[0], hx = in[1], hy = in[2], a = in[3], ax = in[4], ay = in[5], nov = in[6], nol = in[7],
            c = in[8], eta = in[9], u1 = in[10], u2 = in[11];
    const V3<R> wo(wo3[0], wo3[1], wo3[2]), n(n3[0], n3[1], n3[2]), F0(f0[0], f0[1], f0[2]);
    out[0] = D_GGX(noh, a);
    out[1] = D_GGX_Aniso(noh, hx, hy, ax, ay);
    out[2] = G1_Smith(nov, a);
    out[3] = G2_Smith(nol, nov, a);
    out[4] = F_SchlickS(c, R(0.04));
    V3<R> fs3 = F_SchlickV(c, F0);
    out[5] = fs3.x; out[6] = fs3.y; out[7] = fs3.z;
    out[8] = F_Dielectric(c, eta);
    V3<R> h = sampleVNDF(wo, ax, ay, u1, u2);
    out[9] = h.x; out[10] = h.y; out[11] = h.z;
    out[12] = pdfVNDF(wo, h, ax, ay);
    V3<R> wt(0);
    out[13] = refractDir(wo, V3<R>(0, 0, 1), eta, wt) ? 1.0f : 0.0f;
    out[14] = wt.x; out[15] = wt.y; out[16] = wt.z;
    V3<R> cs = sampleCosineHemisphere(u1, u2);
    out[17] = cs.x; out[18] = cs.y; out[19] = cs.z;
    out[20] = powerHeuristic(nov, nol);
    Frame<R> fr(n);
    V3<R> l = fr.toLocal(wo), w = fr.toWorld(wo);
    out[21] = l.x; out[22] = l.y; out[23] = l.z;
    out[24] = w.x; out[25] = w.y; out[26] = w.z;
}

}  // extern "C"

namespace {

template <class Fn>
void parallelFor(size_t n, int nt, Fn&& fn) {
    std::atomic<size_t> next{0};
    auto work = [&]() {
        while (true) {
            size_t i = next.fetch_add(1);
            if (i >= n) break;
            fn(i);
        }
    };
    std::vector<std::thread> pool;
    for (int i = 1; i < nt; ++i) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
}

int threadCount(int threads) {
    return threads > 0 ? threads : static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
}

bool weighted(const float* w, size_t px) { return w[3 * px] != 0.0f || w[3 * px + 1] != 0.0f || w[3 * px + 2] != 0.0f; }

// The tape of every path of the batch (kernelPathTrace's sampler seeding,
// path.cu:216-228) at pixels whose weights are not all zero.
std::vector<std::vector<TapeVertex>> recordTapes(const oracle_scene* s, const pt_camera* cam, int spf, int maxDepth,
                                                 int sampler, uint32_t seed, uint32_t frame, const float* weights,
                                                 int nt) {
    pt_camera_frame cf;
    oracle_camera_frame(cam, &cf);
    const int W = cam->width, H = cam->height;
    std::vector<std::vector<TapeVertex>> tapes(static_cast<size_t>(W) * H * spf);
    parallelFor(static_cast<size_t>(W) * H, nt, [&](size_t px) {
        if (!weighted(weights, px)) return;
        const int x = static_cast<int>(px % W), y = static_cast<int>(px / W);
        Ctx c;
        for (int i = 0; i < spf; ++i) {
            Sampler rng;
            if (sampler == 1) rng.initSobol(x, y, frame * static_cast<uint32_t>(spf) + static_cast<uint32_t>(i), seed);
            else rng.initPCG(x, y, frame, static_cast<uint32_t>(i), seed);
            tapes[px * spf + i] = pathTape(s, cf, W, H, maxDepth, x, y, rng, c);
        }
    });
    return tapes;
}

}  // namespace

#ifndef ORACLE_CXX_FLAGS
#define ORACLE_CXX_FLAGS "unknown"
#endif

extern "C" {

const char* oracle_build_flags() { return ORACLE_CXX_FLAGS; }

void oracle_adjoint_fd(const oracle_scene* s, const pt_camera* cam, int spf, int maxDepth, int sampler, uint32_t seed,
                       uint32_t frame, const float* weights, int nparams, const int* kind, const int* pa, const int* pb,
                       const int* pc, double h, int threads, double* grad, double* value) {
    const int W = cam->width, H = cam->height;
    const int nt = threadCount(threads);
    const auto tapes = recordTapes(s, cam, spf, maxDepth, sampler, seed, frame, weights, nt);
    auto image_dot = [&](const oracle_scene* sp) {
        std::vector<double> part(static_cast<size_t>(W) * H, 0.0);
        parallelFor(static_cast<size_t>(W) * H, nt, [&](size_t px) {
            if (!weighted(weights, px)) return;
            double acc = 0.0;
            for (int i = 0; i < spf; ++i) {
                V3<double> L = tapeEval(sp, tapes[px * spf + i]);
                acc += weights[3 * px + 0] * std::fmin(L.x, 100.0) + weights[3 * px + 1] * std::fmin(L.y, 100.0) +
                       weights[3 * px + 2] * std::fmin(L.z, 100.0);
            }
            part[px] = acc / spf;
        });
        double sum = 0.0;
        for (double v : part) sum += v;
        return sum;
    };
    if (value) *value = image_dot(s);
    for (int k = 0; k < nparams; ++k) {
        oracle_scene sp = *s, sm = *s;
        // the parameter is a float: the step actually taken is the difference
        // of the two rounded values
        auto field = [&](oracle_scene& t) -> float& {
            const int a = pa[k], b = pb[k], ch = pc[k];
            switch (kind[k]) {
                case 0: return (&t.m[a].base_color.x)[ch];
                case 1: return t.m[a].roughness;
                case 2: return t.m[a].metallic;
                case 3: return t.tex[a].rgba[4 * static_cast<size_t>(b) + ch];
                default: return t.env.rgba[4 * static_cast<size_t>(b) + ch];
            }
        };
        float& fp = field(sp);
        float& fm = field(sm);
        const double x0 = fp;
        fp = static_cast<float>(x0 + h);
        fm = static_cast<float>(x0 - h);
        grad[k] = (image_dot(&sp) - image_dot(&sm)) / (static_cast<double>(fp) - static_cast<double>(fm));
    }
}

void oracle_adjoint(const oracle_scene* s, const pt_camera* cam, int spf, int maxDepth, int sampler, uint32_t seed,
                    uint32_t frame, const float* weights, int threads, double* g_mat, double* g_tex, double* g_env,
                    double* value) {
    const int W = cam->width, H = cam->height;
    const int nt = threadCount(threads);
    const auto tapes = recordTapes(s, cam, spf, maxDepth, sampler, seed, frame, weights, nt);
    ADTape proto;
    proto.mats = s->m.data();
    proto.nmat = s->m.size();
    int64_t base = static_cast<int64_t>(5 * s->m.size());
    const int64_t tex0 = base;
    for (const Texture& t : s->tex) {
        proto.blocks.push_back({t.rgba.data(), t.rgba.size(), base});
        base += static_cast<int64_t>(t.rgba.size());
    }
    const int64_t env0 = base;
    if (s->env.w > 0) {
        proto.blocks.push_back({s->env.rgba.data(), s->env.rgba.size(), base});
        base += static_cast<int64_t>(s->env.rgba.size());
    }
    proto.nparams = base;
    // per-thread gradients, summed in thread order at the end
    std::vector<std::vector<double>> part(nt);
    std::vector<double> pxval(static_cast<size_t>(W) * H, 0.0);
    std::atomic<size_t> next{0};
    auto work = [&](int ti) {
        ADTape tape = proto;
        std::vector<double>& g = part[ti];
        g.assign(static_cast<size_t>(proto.nparams), 0.0);
        std::vector<double> adj;
        tlTape = &tape;
        while (true) {
            const size_t px = next.fetch_add(1);
            if (px >= static_cast<size_t>(W) * H) break;
            if (!weighted(weights, px)) continue;
            double acc = 0.0;
            for (int i = 0; i < spf; ++i) {
                tape.nodes.clear();
                tape.leaves.clear();
                const V3<AD> L = tapeEvalT<AD>(s, tapes[px * spf + i]);
                adj.assign(tape.nodes.size(), 0.0);
                const AD Lc[3] = {L.x, L.y, L.z};
                for (int c = 0; c < 3; ++c) {
                    const double w = weights[3 * px + c];
                    acc += w * std::fmin(Lc[c].v, 100.0);
                    // the clamp at 100 (path.cu:201-204) passes no derivative above it
                    if (Lc[c].i >= 0 && Lc[c].v < 100.0) adj[Lc[c].i] += w / spf;
                }
                for (size_t k = tape.nodes.size(); k-- > 0;) {
                    const double a = adj[k];
                    if (a == 0.0) continue;
                    const ADNode& n = tape.nodes[k];
                    if (n.a >= 0) adj[n.a] += a * n.da;
                    if (n.b >= 0) adj[n.b] += a * n.db;
                }
                for (const auto& lf : tape.leaves) g[lf.second] += adj[lf.first];
            }
            pxval[px] = acc / spf;
        }
        tlTape = nullptr;
    };
    std::vector<std::thread> pool;
    for (int i = 1; i < nt; ++i) pool.emplace_back(work, i);
    work(0);
    for (auto& t : pool) t.join();
    std::vector<double> g(static_cast<size_t>(proto.nparams), 0.0);
    for (const auto& p : part)
        for (size_t k = 0; k < g.size(); ++k) g[k] += p[k];
    if (g_mat) std::memcpy(g_mat, g.data(), sizeof(double) * 5 * s->m.size());
    if (g_tex) std::memcpy(g_tex, g.data() + tex0, sizeof(double) * static_cast<size_t>(env0 - tex0));
    if (g_env && s->env.w > 0) std::memcpy(g_env, g.data() + env0, sizeof(double) * s->env.rgba.size());
    if (value) {
        double sum = 0.0;
        for (double v : pxval) sum += v;
        *value = sum;
    }
}

}  // extern "C"

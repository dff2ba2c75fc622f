// Binary -> 4-wide BVH collapse and child-box quantization (see bvh_wide.h).
#include "bvh_wide.h"

#include <cmath>
#include <cstring>
#include <deque>

namespace ptamd {

namespace {

constexpr uint32_t kLeaf = 0x80000000u;

struct BinNode {
    float lo[3], hi[3];
    uint32_t d0, d1;
    bool leaf() const { return (d0 & kLeaf) != 0; }
};

BinNode read_node(const float* nodes8, uint32_t k) {
    BinNode n;
    const float* p = nodes8 + 8 * static_cast<size_t>(k);
    for (int a = 0; a < 3; ++a) n.lo[a] = p[a], n.hi[a] = p[4 + a];
    std::memcpy(&n.d0, p + 3, 4);
    std::memcpy(&n.d1, p + 7, 4);
    return n;
}

float area(const BinNode& n) {
    const float dx = n.hi[0] - n.lo[0], dy = n.hi[1] - n.lo[1], dz = n.hi[2] - n.lo[2];
    return 2.0f * (dx * dy + dx * dz + dy * dz);
}

// The decode the traversal kernels use: fma(q, s, p) in single precision
// (q * s is exact: an integer below 256 times a power of two).
inline float decode(uint32_t q, float s, float p) { return std::fmaf(static_cast<float>(q), s, p); }

// Scale for one axis: the smallest power of two s with p + 255 s >= q (decoded).
float axis_scale(float p, float q) {
    const float e = q - p;
    if (!(e > 0.0f)) return 1.0f;  // flat axis: every byte 0 decodes to p
    int ex = 0;
    std::frexp(e / 255.0f, &ex);   // e / 255 < 2^ex
    float s = std::ldexp(1.0f, ex - 1);
    while (decode(255, s, p) < q) s *= 2.0f;
    while (s > 0.0f && decode(255, s * 0.5f, p) >= q && s * 0.5f >= std::ldexp(1.0f, -126)) s *= 0.5f;
    return s;
}

uint32_t quant_lo(float v, float s, float p) {
    double g = std::floor((static_cast<double>(v) - p) / s);
    uint32_t q = g <= 0.0 ? 0u : g >= 255.0 ? 255u : static_cast<uint32_t>(g);
    while (q > 0 && decode(q, s, p) > v) --q;  // q = 0 decodes to p <= v
    return q;
}

uint32_t quant_hi(float v, float s, float p) {
    double g = std::ceil((static_cast<double>(v) - p) / s);
    uint32_t q = g <= 0.0 ? 0u : g >= 255.0 ? 255u : static_cast<uint32_t>(g);
    while (q < 255 && decode(q, s, p) < v) ++q;  // q = 255 decodes to >= the node's top
    return q;
}

}  // namespace

bool collapse_wide_bvh(const float* nodes8, uint32_t num_nodes, WideBvh* out, std::string* err) {
    *out = WideBvh();
    if (num_nodes == 0) return true;
    auto fail = [&](const char* m) {
        if (err) *err = m;
        *out = WideBvh();
        return false;
    };
    struct Item {
        uint32_t bin;    // binary node this wide node expands
        uint32_t stack;  // stack entries its ancestors can leave pushed
        uint32_t level;
    };
    std::deque<Item> todo;
    todo.push_back({0u, 0u, 1u});
    out->num_nodes = 1;
    std::vector<uint32_t> visits(num_nodes, 0);
    size_t w = 0;
    while (!todo.empty()) {
        const Item it = todo.front();
        todo.pop_front();
        const BinNode b = read_node(nodes8, it.bin);
        uint32_t set[4];
        int n = 0;
        if (b.leaf()) {
            set[n++] = it.bin;  // a leaf root: one leaf child
        } else {
            if (b.d0 >= num_nodes || b.d1 >= num_nodes) return fail("BVH child index out of range");
            set[n++] = b.d0;
            set[n++] = b.d1;
        }
        while (n < 4) {  // open the internal child with the largest surface area
            int best = -1;
            float best_a = -1.0f;
            for (int i = 0; i < n; ++i) {
                const BinNode c = read_node(nodes8, set[i]);
                if (!c.leaf() && area(c) > best_a) best = i, best_a = area(c);
            }
            if (best < 0) break;
            const BinNode c = read_node(nodes8, set[best]);
            if (c.d0 >= num_nodes || c.d1 >= num_nodes) return fail("BVH child index out of range");
            for (int i = n; i > best + 1; --i) set[i] = set[i - 1];  // keep the spatial order
            set[best] = c.d0;
            set[best + 1] = c.d1;
            ++n;
        }
        BinNode kids[4];
        float P[3], Q[3], S[3];
        for (int a = 0; a < 3; ++a) P[a] = INFINITY, Q[a] = -INFINITY;
        for (int i = 0; i < n; ++i) {
            if (++visits[set[i]] > 1) return fail("BVH node reached twice (not a tree)");
            kids[i] = read_node(nodes8, set[i]);
            for (int a = 0; a < 3; ++a) {
                if (!std::isfinite(kids[i].lo[a]) || !std::isfinite(kids[i].hi[a]))
                    return fail("non-finite BVH bounds");
                P[a] = std::fmin(P[a], kids[i].lo[a]), Q[a] = std::fmax(Q[a], kids[i].hi[a]);
            }
        }
        for (int a = 0; a < 3; ++a) S[a] = axis_scale(P[a], Q[a]);
        uint32_t qlo[3] = {0, 0, 0}, qhi[3] = {0, 0, 0}, words[4] = {kWideEmpty, kWideEmpty, kWideEmpty, kWideEmpty};
        const uint32_t pushed = it.stack + static_cast<uint32_t>(n - 1);
        out->max_stack = std::max(out->max_stack, pushed);
        out->depth = std::max(out->depth, it.level);
        for (int i = 0; i < n; ++i) {
            for (int a = 0; a < 3; ++a) {
                qlo[a] |= quant_lo(kids[i].lo[a], S[a], P[a]) << (8 * i);
                qhi[a] |= quant_hi(kids[i].hi[a], S[a], P[a]) << (8 * i);
            }
            if (kids[i].leaf()) {
                words[i] = kids[i].d0;
            } else {
                words[i] = out->num_nodes++;
                todo.push_back({set[i], pushed, it.level + 1});
            }
        }
        // unused slots: bytes 0 (their word marks them empty)
        out->words.resize(16 * (w + 1));
        float* o = &out->words[16 * w];
        o[0] = P[0], o[1] = P[1], o[2] = P[2], o[3] = S[0];
        o[4] = S[1], o[5] = S[2];
        std::memcpy(o + 6, &qlo[0], 4);
        std::memcpy(o + 7, &qlo[1], 4);
        std::memcpy(o + 8, &qlo[2], 4);
        std::memcpy(o + 9, &qhi[0], 4);
        std::memcpy(o + 10, &qhi[1], 4);
        std::memcpy(o + 11, &qhi[2], 4);
        std::memcpy(o + 12, words, 16);
        ++w;
    }
    if (w != out->num_nodes) return fail("wide BVH node count mismatch");
    return true;
}

}  // namespace ptamd

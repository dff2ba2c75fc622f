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


// Which binary nodes become wide nodes.  A wide node costs one fetch and
// four box tests whenever a ray enters its box, so the expected traversal
// cost of a collapse is proportional to the summed surface area of the
// binary nodes chosen as wide nodes (surface area heuristic); every binary
// leaf ends up in exactly one wide node whatever the choice, so the leaves'
// cost is fixed.  best[k] = the least cost of covering a subtree by at most
// k children of its parent wide node (k = 1: the subtree's root is one child,
// a leaf or a wide node of its own), computed bottom-up:
//   wide(n)   = A(n) + min_j best_l[j] + best_r[4 - j]   (n opened as a wide node)
//   best_n[1] = leaf ? 0 : wide(n)
//   best_n[k] = min(best_n[1], min_j best_l[j] + best_r[k - j])
// split[k] records the j of the minimum (0: n stays one child).  This is the
// dynamic program of Ylitie, Karras and Laine, "Efficient incoherent ray
// traversal on GPUs through compressed wide BVHs" (HPG 2017), restricted to
// binary leaves kept as they are.
struct CutPlan {
    std::vector<double> best;     // 5 per node (index 1..4)
    std::vector<uint8_t> split;   // 5 per node
    std::vector<uint8_t> split4;  // j of wide(n)
};

bool plan_cuts(const float* nodes8, uint32_t num_nodes, CutPlan* plan, std::string* err) {
    plan->best.assign(5 * static_cast<size_t>(num_nodes), 0.0);
    plan->split.assign(5 * static_cast<size_t>(num_nodes), 0);
    plan->split4.assign(num_nodes, 1);
    // breadth-first order from the root: children after parents
    std::vector<uint32_t> order;
    std::vector<uint8_t> seen(num_nodes, 0);
    order.push_back(0);
    seen[0] = 1;
    for (size_t i = 0; i < order.size(); ++i) {
        const BinNode b = read_node(nodes8, order[i]);
        if (b.leaf()) continue;
        for (uint32_t c : {b.d0, b.d1}) {
            if (c >= num_nodes) {
                if (err) *err = "BVH child index out of range";
                return false;
            }
            if (seen[c]) {
                if (err) *err = "BVH node reached twice (not a tree)";
                return false;
            }
            seen[c] = 1;
            order.push_back(c);
        }
    }
    for (size_t i = order.size(); i-- > 0;) {
        const uint32_t k = order[i];
        const BinNode b = read_node(nodes8, k);
        double* bn = &plan->best[5 * static_cast<size_t>(k)];
        uint8_t* sp = &plan->split[5 * static_cast<size_t>(k)];
        if (b.leaf()) {
            for (int m = 1; m <= 4; ++m) bn[m] = 0.0, sp[m] = 0;
            continue;
        }
        const double* bl = &plan->best[5 * static_cast<size_t>(b.d0)];
        const double* br = &plan->best[5 * static_cast<size_t>(b.d1)];
        auto pair_min = [&](int m, int* arg) {
            double v = INFINITY;
            for (int j = 1; j < m; ++j)
                if (bl[j] + br[m - j] < v) v = bl[j] + br[m - j], *arg = j;
            return v;
        };
        int j4 = 1;
        bn[1] = static_cast<double>(area(b)) + pair_min(4, &j4);
        sp[1] = 0;
        plan->split4[k] = static_cast<uint8_t>(j4);
        for (int m = 2; m <= 4; ++m) {
            int j = 1;
            const double v = pair_min(m, &j);
            if (v < bn[1]) bn[m] = v, sp[m] = static_cast<uint8_t>(j);
            else bn[m] = bn[1], sp[m] = 0;
        }
    }
    return true;
}

// Appends the children that subtree n contributes with a budget of m slots.
void emit_cut(const CutPlan& plan, const float* nodes8, uint32_t n, int m, uint32_t* set, int* count) {
    const int j = plan.split[5 * static_cast<size_t>(n) + m];
    if (j == 0) {
        set[(*count)++] = n;
        return;
    }
    const BinNode b = read_node(nodes8, n);
    emit_cut(plan, nodes8, b.d0, j, set, count);
    emit_cut(plan, nodes8, b.d1, m - j, set, count);
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
    CutPlan plan;
    if (!plan_cuts(nodes8, num_nodes, &plan, err)) {
        *out = WideBvh();
        return false;
    }
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
        } else {  // the cut the cost table chose: up to four descendants, in spatial order
            emit_cut(plan, nodes8, b.d0, plan.split4[it.bin], set, &n);
            emit_cut(plan, nodes8, b.d1, 4 - plan.split4[it.bin], set, &n);
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
                // on equal values (+0 and -0) the later child's value is kept
                P[a] = P[a] < kids[i].lo[a] ? P[a] : kids[i].lo[a];
                Q[a] = Q[a] > kids[i].hi[a] ? Q[a] : kids[i].hi[a];
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

// Binary -> 4-wide BVH collapse on the GPU (see bvh_wide_build.h).
//
// Two sweeps over the binary tree the level-synchronous builder made:
//   1. bottom-up, one launch per binary level (deepest first): the surface-
//      area program of the host collapse (host/bvh_wide.cpp): per node the
//      least cost best[m] of covering its subtree with m = 1..4 children of
//      the parent wide node, and the split that reaches it;
//   2. top-down, one launch pair per wide level: every wide node of the level
//      expands its binary node into the cut the program chose (up to four
//      descendants, left to right), a prefix sum over the level numbers the
//      children that are wide nodes themselves, and the node is written with
//      its quantized child boxes.  Numbering a level's children by (parent,
//      slot) is the host collapse's breadth-first queue order, so the nodes,
//      words and boxes equal the host's bit for bit.
// Contraction is off so the float area and decode sequences round as the host
// code's do.
#pragma clang fp contract(off)

#include "bvh_wide_build.h"

#include <algorithm>

#include "bvh_build.h"

namespace ptd {

namespace {

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr int kBlock = 256;
constexpr uint32_t kMaxWideLevels = 256;

// best[1..4] of one binary node.
struct Cost {
    double b[4];
};

// plan word of a binary node: bits 0-3 the split of the node opened as a wide
// node (children of the left subtree), bits 4(m-1)..4(m-1)+3 for m = 2..4 the
// split of a budget of m slots (0: the node stays one child).
__device__ __forceinline__ uint32_t plan_split(uint32_t plan, int m) {
    return m <= 1 ? 0u : (plan >> (4 * (m - 1))) & 15u;
}

__device__ __forceinline__ uint32_t word0(const float4* nodes, uint32_t k) { return __float_as_uint(nodes[2 * k].w); }
__device__ __forceinline__ uint32_t word1(const float4* nodes, uint32_t k) {
    return __float_as_uint(nodes[2 * k + 1].w);
}

__global__ __launch_bounds__(kBlock) void k_wide_cost(const float4* __restrict__ nodes, uint32_t lb, uint32_t K,
                                                      Cost* best, uint32_t* plan) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= K) return;
    const uint32_t k = lb + t;
    const float4 a = nodes[2 * k], b = nodes[2 * k + 1];
    const uint32_t d0 = __float_as_uint(a.w), d1 = __float_as_uint(b.w);
    if (d0 & kLeafBit) {
        best[k] = Cost{{0.0, 0.0, 0.0, 0.0}};
        plan[k] = 1u;
        return;
    }
    const Cost L = best[d0], R = best[d1];
    const float dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
    const float area = 2.0f * (dx * dy + dx * dz + dy * dz);
    // min over j of L[j] + R[m - j], the first j on ties
    auto pair_min = [&](int m, int* arg) {
        double v = __builtin_inf();
        for (int j = 1; j < m; ++j) {
            const double c = L.b[j - 1] + R.b[m - j - 1];
            if (c < v) v = c, *arg = j;
        }
        return v;
    };
    Cost out;
    int j4 = 1;
    out.b[0] = static_cast<double>(area) + pair_min(4, &j4);
    uint32_t p = static_cast<uint32_t>(j4);
    for (int m = 2; m <= 4; ++m) {
        int j = 1;
        const double v = pair_min(m, &j);
        if (v < out.b[0]) {
            out.b[m - 1] = v;
            p |= static_cast<uint32_t>(j) << (4 * (m - 1));
        } else {
            out.b[m - 1] = out.b[0];
        }
    }
    best[k] = out;
    plan[k] = p;
}

// The up to four binary nodes wide node `bin` gets as children, left to right.
__device__ int cut_set(const float4* __restrict__ nodes, const uint32_t* __restrict__ plan, uint32_t bin,
                       uint32_t set[4]) {
    const uint32_t d0 = word0(nodes, bin);
    if (d0 & kLeafBit) {
        set[0] = bin;  // a leaf root: one leaf child
        return 1;
    }
    const uint32_t j4 = plan[bin] & 15u;
    // pending (node, budget) pairs, the next one on top; at most four, since
    // the budgets of the pending pairs and the emitted children sum to 4
    uint32_t sn[4], sm[4];
    int sp = 0, n = 0;
    sn[sp] = word1(nodes, bin), sm[sp++] = 4u - j4;
    sn[sp] = d0, sm[sp++] = j4;
    while (sp > 0 && n < 4) {
        --sp;
        const uint32_t k = sn[sp], m = sm[sp];
        const uint32_t j = plan_split(plan[k], static_cast<int>(m));
        if (j == 0 || sp + 2 > 4) {
            set[n++] = k;
        } else {
            sn[sp] = word1(nodes, k), sm[sp++] = m - j;
            sn[sp] = word0(nodes, k), sm[sp++] = j;
        }
    }
    return n;
}

// The decode the traversal kernels use (an integer below 256 times a power of
// two, plus the corner, in single precision).
__device__ __forceinline__ float decode(uint32_t q, float s, float p) { return fmaf(static_cast<float>(q), s, p); }

// The smallest power of two s with decode(255, s, p) >= q.
__device__ float axis_scale(float p, float q) {
    const float e = q - p;
    if (!(e > 0.0f)) return 1.0f;
    int ex = 0;
    (void)frexpf(e / 255.0f, &ex);
    float s = ldexpf(1.0f, ex - 1);
    while (decode(255, s, p) < q) s *= 2.0f;
    while (s > 0.0f && decode(255, s * 0.5f, p) >= q && s * 0.5f >= ldexpf(1.0f, -126)) s *= 0.5f;
    return s;
}

__device__ uint32_t quant_lo(float v, float s, float p) {
    const double g = floor((static_cast<double>(v) - p) / s);
    uint32_t q = g <= 0.0 ? 0u : g >= 255.0 ? 255u : static_cast<uint32_t>(g);
    while (q > 0 && decode(q, s, p) > v) --q;
    return q;
}

__device__ uint32_t quant_hi(float v, float s, float p) {
    const double g = ceil((static_cast<double>(v) - p) / s);
    uint32_t q = g <= 0.0 ? 0u : g >= 255.0 ? 255u : static_cast<uint32_t>(g);
    while (q < 255 && decode(q, s, p) < v) ++q;
    return q;
}

__global__ void k_wide_root(uint2* items) { items[0] = make_uint2(0u, 0u); }

// Pass 1 of a wide level: each item's cut and how many of its children are
// wide nodes.
__global__ __launch_bounds__(kBlock) void k_wide_count(const float4* __restrict__ nodes,
                                                       const uint32_t* __restrict__ plan,
                                                       const uint2* __restrict__ items, uint32_t K, uint4* sets,
                                                       uint32_t* nset, uint32_t* inner) {
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= K) return;
    uint32_t set[4] = {0u, 0u, 0u, 0u};
    const int n = cut_set(nodes, plan, items[j].x, set);
    uint32_t c = 0;
    for (int i = 0; i < n; ++i) c += (word0(nodes, set[i]) & kLeafBit) ? 0u : 1u;
    sets[j] = make_uint4(set[0], set[1], set[2], set[3]);
    nset[j] = static_cast<uint32_t>(n);
    inner[j] = c;
}

// Pass 2: the level's nodes and the next level's items.  status[0]: bit 0
// non-finite bounds, bit 1 capacity exceeded; status[1]: the deepest stack.
__global__ __launch_bounds__(kBlock) void k_wide_emit(const float4* __restrict__ nodes,
                                                      const uint2* __restrict__ items, uint32_t K,
                                                      const uint4* __restrict__ sets, const uint32_t* __restrict__ nset,
                                                      const uint32_t* __restrict__ offs, uint32_t base,
                                                      uint32_t cap, float4* wide, uint2* next, uint32_t* status) {
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= K) return;
    const uint4 s4 = sets[j];
    const uint32_t set[4] = {s4.x, s4.y, s4.z, s4.w};
    const int n = static_cast<int>(nset[j]);
    float lo[4][3], hi[4][3];
    uint32_t w0[4];
    float P[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
    float Q[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
    bool finite = true;
    for (int i = 0; i < n; ++i) {
        const float4 a = nodes[2 * set[i]], b = nodes[2 * set[i] + 1];
        lo[i][0] = a.x, lo[i][1] = a.y, lo[i][2] = a.z;
        hi[i][0] = b.x, hi[i][1] = b.y, hi[i][2] = b.z;
        w0[i] = __float_as_uint(a.w);
        for (int ax = 0; ax < 3; ++ax) {
            finite = finite && isfinite(lo[i][ax]) && isfinite(hi[i][ax]);
            P[ax] = P[ax] < lo[i][ax] ? P[ax] : lo[i][ax];
            Q[ax] = Q[ax] > hi[i][ax] ? Q[ax] : hi[i][ax];
        }
    }
    // a non-finite box voids the tree; the node and its children are still
    // written so that the next levels read only indices of this tree
    if (!finite) atomicOr(status, 1u);
    const uint32_t out = base + j;
    if (out >= cap) {
        atomicOr(status, 2u);
        return;
    }
    const float S[3] = {axis_scale(P[0], Q[0]), axis_scale(P[1], Q[1]), axis_scale(P[2], Q[2])};
    const uint32_t pushed = items[j].y + static_cast<uint32_t>(n - 1);
    atomicMax(status + 1, pushed);
    uint32_t qlo[3] = {0u, 0u, 0u}, qhi[3] = {0u, 0u, 0u};
    uint32_t words[4] = {kEmpty, kEmpty, kEmpty, kEmpty};
    uint32_t r = 0;
    const uint32_t next_base = base + K;  // the first index of the next level
    for (int i = 0; i < n; ++i) {
        for (int ax = 0; ax < 3; ++ax) {
            qlo[ax] |= quant_lo(lo[i][ax], S[ax], P[ax]) << (8 * i);
            qhi[ax] |= quant_hi(hi[i][ax], S[ax], P[ax]) << (8 * i);
        }
        if (w0[i] & kLeafBit) {
            words[i] = w0[i];
        } else {
            const uint32_t slot = offs[j] + r++;
            if (next_base + slot >= cap) {
                atomicOr(status, 2u);
                return;
            }
            words[i] = next_base + slot;
            next[slot] = make_uint2(set[i], pushed);
        }
    }
    float4* o = wide + 4 * static_cast<size_t>(out);
    o[0] = make_float4(P[0], P[1], P[2], S[0]);
    o[1] = make_float4(S[1], S[2], __uint_as_float(qlo[0]), __uint_as_float(qlo[1]));
    o[2] = make_float4(__uint_as_float(qlo[2]), __uint_as_float(qhi[0]), __uint_as_float(qhi[1]),
                       __uint_as_float(qhi[2]));
    o[3] = make_float4(__uint_as_float(words[0]), __uint_as_float(words[1]), __uint_as_float(words[2]),
                       __uint_as_float(words[3]));
}

inline uint32_t nblocks(size_t n) { return static_cast<uint32_t>((n + kBlock - 1) / kBlock); }

}  // namespace

bool gpu_collapse_wide(const float4* nodes, uint32_t num_nodes, const std::vector<uint32_t>& level_begin,
                       const std::vector<uint32_t>& level_count, hipStream_t stream, WideDeviceResult* out,
                       std::string* err) {
    *out = WideDeviceResult();
    if (num_nodes == 0) return true;
    // every wide node but a leaf root expands a distinct inner binary node
    // (slot 1 is unused): at most num_nodes / 2 + 1 of them
    const uint32_t cap = num_nodes / 2 + 2;
    Cost* best = nullptr;
    uint32_t *plan = nullptr, *nset = nullptr, *inner = nullptr, *offs = nullptr, *tmp = nullptr, *scal = nullptr;
    uint2 *cur = nullptr, *next = nullptr;
    uint4* sets = nullptr;
    float4* wide = nullptr;
    hipError_t e = hipSuccess;
    auto ok = [&](hipError_t r) {
        if (e == hipSuccess) e = r;
        return e == hipSuccess;
    };
    auto fail = [&](const std::string& m) {
        if (err && err->empty()) *err = m;
        return false;
    };
    bool good = true;
    uint32_t status[2] = {0u, 0u};
    uint32_t base = 0, K = 1, levels = 0;
    if (!(ok(hipMalloc(reinterpret_cast<void**>(&best), sizeof(Cost) * num_nodes)) &&
          ok(hipMalloc(reinterpret_cast<void**>(&plan), sizeof(uint32_t) * num_nodes)) &&
          ok(hipMalloc(reinterpret_cast<void**>(&cur), sizeof(uint2) * cap)) &&
          ok(hipMalloc(reinterpret_cast<void**>(&next), sizeof(uint2) * cap)) &&
          ok(hipMalloc(reinterpret_cast<void**>(&sets), sizeof(uint4) * cap)) &&
          ok(hipMalloc(reinterpret_cast<void**>(&nset), sizeof(uint32_t) * cap)) &&
          ok(hipMalloc(reinterpret_cast<void**>(&inner), sizeof(uint32_t) * cap)) &&
          ok(hipMalloc(reinterpret_cast<void**>(&offs), sizeof(uint32_t) * cap)) &&
          ok(hipMalloc(reinterpret_cast<void**>(&tmp), sizeof(uint32_t) * (2048 + 32))) &&
          ok(hipMalloc(reinterpret_cast<void**>(&scal), sizeof(uint32_t) * 4)) &&
          ok(hipMalloc(reinterpret_cast<void**>(&wide), sizeof(float4) * 4 * static_cast<size_t>(cap))))) {
        good = fail(std::string("wide BVH buffers: ") + hipGetErrorString(e));
    }
    if (good) {
        // 1. the surface-area program, deepest binary level first
        for (size_t l = level_begin.size(); l-- > 0;) {
            const uint32_t lb = level_begin[l], lk = level_count[l];
            if (lk == 0) continue;
            if (static_cast<uint64_t>(lb) + lk > num_nodes) {
                good = fail("wide BVH: binary level outside the node array");
                break;
            }
            hipLaunchKernelGGL(k_wide_cost, dim3(nblocks(lk)), dim3(kBlock), 0, stream, nodes, lb, lk, best, plan);
        }
    }
    if (good) {
        // 2. the wide levels, root first
        hipLaunchKernelGGL(k_wide_root, dim3(1), dim3(1), 0, stream, cur);
        good = ok(hipMemsetAsync(scal, 0, sizeof(uint32_t) * 4, stream)) || fail("wide BVH: clear status");
        while (good && K > 0) {
            if (++levels > kMaxWideLevels || static_cast<uint64_t>(base) + K > cap) {
                good = fail("wide BVH: not a tree (too many wide nodes or levels)");
                break;
            }
            hipLaunchKernelGGL(k_wide_count, dim3(nblocks(K)), dim3(kBlock), 0, stream, nodes, plan, cur, K, sets,
                               nset, inner);
            scan_exclusive_u32(inner, offs, K, tmp, scal + 2, stream);
            hipLaunchKernelGGL(k_wide_emit, dim3(nblocks(K)), dim3(kBlock), 0, stream, nodes, cur, K, sets, nset,
                               offs, base, cap, wide, next, scal);
            uint32_t total = 0;
            if (!(ok(hipGetLastError()) &&
                  ok(hipMemcpyAsync(&total, scal + 2, sizeof(uint32_t), hipMemcpyDeviceToHost, stream)) &&
                  ok(hipStreamSynchronize(stream)))) {
                good = fail(std::string("wide BVH level: ") + hipGetErrorString(e));
                break;
            }
            base += K;
            K = total;
            std::swap(cur, next);
        }
    }
    if (good && ok(hipMemcpyAsync(status, scal, sizeof(status), hipMemcpyDeviceToHost, stream)) &&
        ok(hipStreamSynchronize(stream))) {
        if (status[0] & 2u) good = fail("wide BVH: not a tree (node capacity exceeded)");
    } else if (good) {
        good = fail(std::string("wide BVH status: ") + hipGetErrorString(e));
    }
    (void)hipStreamSynchronize(stream);
    for (void* p : {static_cast<void*>(best), static_cast<void*>(plan), static_cast<void*>(cur),
                    static_cast<void*>(next), static_cast<void*>(sets), static_cast<void*>(nset),
                    static_cast<void*>(inner), static_cast<void*>(offs), static_cast<void*>(tmp),
                    static_cast<void*>(scal)})
        if (p) (void)hipFree(p);
    if (!good || (status[0] & 1u)) {
        if (wide) (void)hipFree(wide);
        if (good) out->finite = false;
        return good;
    }
    out->nodes = wide;
    out->num_nodes = base;
    out->max_stack = status[1];
    out->depth = levels;
    return true;
}

}  // namespace ptd

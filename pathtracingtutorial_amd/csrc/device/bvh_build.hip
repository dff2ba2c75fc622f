// GPU binned-SAH BVH build for gfx950.
//
// Level-synchronous top-down build.  Every level runs over all primitive
// slots (one thread per slot, primitives of one node are contiguous):
//   1. centroid bounds of every node of the level that may split,
//   2. binning of centroids into B buckets (count + AABB per bucket),
//   3. one thread per node sweeps the B-1 split planes,
//   4. children are allocated adjacently (left, left+1) by a prefix sum over
//      the split flags, so the final array needs no reordering for the
//      sibling-pair traversal layout,
//   5. a stable partition (prefix sum of "goes left" flags) scatters the
//      primitives into the children's ranges.
// Blocks whose 256 slots all belong to one node reduce in LDS and issue one
// atomic per value; mixed blocks use global atomics directly.
//
// The split decision follows the reference builder arithmetic exactly
// (Part_2/src/accel/bvh_builder.cpp:92-177): max-extent axis of the centroid
// bounds, bucket = int(B * (c - lo) / extent) clamped to B-1, forward/backward
// sweeps with float costs count*SA, cost = 0.5 + min/SA(node), leaf if count
// < leaf_threshold or count < cost or all primitives fall left.  With B = 24
// and leaf_threshold = 4 the tree is the reference's tree node for node (only
// the order of triangles inside a leaf may differ).  Contraction into FMA is
// disabled in this file so the float sequence matches.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "bvh_build.h"
#include "dev_scene.h"

namespace ptd {

namespace {

constexpr uint32_t kDone = 0xFFFFFFFFu;
constexpr float kFltMax = 3.402823466e+38f;
constexpr int kBlock = 256;
constexpr int kMaxBins = 32;
constexpr int kScanItems = 8;
constexpr int kScanBlock = kBlock * kScanItems;  // 2048

// Order-preserving float <-> uint map for atomic min/max.
__device__ __forceinline__ uint32_t fenc(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float fdec(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u ^ 0x80000000u) : ~u);
}
static uint32_t fenc_host(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// accel/bvh_builder.cpp:35-41
__device__ __forceinline__ float sa6(const float lo[3], const float hi[3]) {
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    if (dx < 0 || dy < 0 || dz < 0) return 0.0f;
    return 2.0f * (dx * dy + dx * dz + dy * dz);
}
// accel/bvh_builder.cpp:43-49
__device__ __forceinline__ int max_dim(const float lo[3], const float hi[3]) {
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    if (dx > dy && dx > dz) return 0;
    return dy > dz ? 1 : 2;
}

struct Level {
    uint32_t lb, K;  // nodes [lb, lb+K)
};

// Per-node split information of the current level.
struct SplitInfo {
    int32_t mode;       // 0 leaf, 1 bucket split, 2 median split
    int32_t axis;
    int32_t bucket;     // last bucket that goes left
    uint32_t left_count;
};

__device__ __forceinline__ bool slot_node(const float4* prims, uint32_t i, uint32_t& node) {
    node = __float_as_uint(prims[2 * i + 1].w);
    return node != kDone;
}

__global__ void k_prim_init(const float4* __restrict__ vpos, const uint4* __restrict__ faces, uint32_t n,
                            float4* __restrict__ prims) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint4 f = faces[i];
    float4 a = vpos[f.x], b = vpos[f.y], c = vpos[f.z];
    // accel/bvh_builder.cpp:62-75
    float4 lo = make_float4(fminf(fminf(a.x, b.x), c.x), fminf(fminf(a.y, b.y), c.y), fminf(fminf(a.z, b.z), c.z),
                            __uint_as_float(i));
    float4 hi = make_float4(fmaxf(fmaxf(a.x, b.x), c.x), fmaxf(fmaxf(a.y, b.y), c.y), fmaxf(fmaxf(a.z, b.z), c.z),
                            __uint_as_float(0u));
    prims[2 * i] = lo;
    prims[2 * i + 1] = hi;
}

__device__ __forceinline__ float wave_min(float v) {
    for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}

// Bounds of the level's nodes: mode 0 = primitive AABBs (root), mode 1 =
// centroids.  Nodes below the leaf threshold are skipped in mode 1.
__global__ void k_node_bounds(const float4* __restrict__ prims, uint32_t n, Level lv, int mode,
                              const uint32_t* __restrict__ ncount, uint32_t thresh, uint32_t* __restrict__ out6) {
    __shared__ float red[6][kBlock / 64];
    __shared__ uint32_t hom_node;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t s0 = blockIdx.x * blockDim.x;
    uint32_t s1 = min(s0 + blockDim.x, n) - 1;
    if (threadIdx.x == 0) {
        uint32_t a, b;
        bool va = slot_node(prims, s0, a), vb = slot_node(prims, s1, b);
        hom_node = (va && vb && a == b) ? a : kDone;
    }
    __syncthreads();
    float lo[3] = {kFltMax, kFltMax, kFltMax}, hi[3] = {-kFltMax, -kFltMax, -kFltMax};
    uint32_t node = kDone;
    bool active = false;
    if (i < n && slot_node(prims, i, node) && node >= lv.lb && node < lv.lb + lv.K &&
        (mode == 0 || ncount[node] >= thresh)) {
        float4 a = prims[2 * i], b = prims[2 * i + 1];
        active = true;
        if (mode == 0) {
            lo[0] = a.x; lo[1] = a.y; lo[2] = a.z;
            hi[0] = b.x; hi[1] = b.y; hi[2] = b.z;
        } else {
            float c0 = 0.5f * (a.x + b.x), c1 = 0.5f * (a.y + b.y), c2 = 0.5f * (a.z + b.z);
            lo[0] = hi[0] = c0; lo[1] = hi[1] = c1; lo[2] = hi[2] = c2;
        }
    }
    uint32_t hn = hom_node;
    bool hom_active = hn != kDone && hn >= lv.lb && hn < lv.lb + lv.K && (mode == 0 || ncount[hn] >= thresh);
    if (hom_active) {
        int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        for (int k = 0; k < 3; ++k) {
            float m0 = wave_min(lo[k]), m1 = wave_max(hi[k]);
            if (lane == 0) { red[k][w] = m0; red[3 + k][w] = m1; }
        }
        __syncthreads();
        if (threadIdx.x < 6) {
            int k = threadIdx.x;
            float v = red[k][0];
            for (int q = 1; q < kBlock / 64; ++q) v = (k < 3) ? fminf(v, red[k][q]) : fmaxf(v, red[k][q]);
            uint32_t* dst = out6 + (hn - lv.lb) * 6 + k;
            if (k < 3) atomicMin(dst, fenc(v));
            else atomicMax(dst, fenc(v));
        }
    } else if (active) {
        uint32_t* dst = out6 + (node - lv.lb) * 6;
        for (int k = 0; k < 3; ++k) {
            atomicMin(dst + k, fenc(lo[k]));
            atomicMax(dst + 3 + k, fenc(hi[k]));
        }
    }
}

// Bucket index of primitive (a, b) for a node with centroid bounds cb (enc).
// accel/bvh_builder.cpp:126-136.
__device__ __forceinline__ int bucket_of(float4 a, float4 b, const uint32_t* cb, int nb, int& axis) {
    float clo[3] = {fdec(cb[0]), fdec(cb[1]), fdec(cb[2])};
    float chi[3] = {fdec(cb[3]), fdec(cb[4]), fdec(cb[5])};
    int dim = max_dim(clo, chi);
    axis = dim;
    float pa = dim == 0 ? a.x : (dim == 1 ? a.y : a.z);
    float pb = dim == 0 ? b.x : (dim == 1 ? b.y : b.z);
    float c = 0.5f * (pa + pb);
    float d = chi[dim] - clo[dim];
    float off = d > 0 ? __fdiv_rn(c - clo[dim], d) : 0.0f;
    int bk = static_cast<int>(static_cast<float>(nb) * off);
    return bk >= nb ? nb - 1 : bk;
}

__global__ void k_bin(const float4* __restrict__ prims, uint32_t n, Level lv, const uint32_t* __restrict__ ncount,
                      uint32_t thresh, const uint32_t* __restrict__ cb, int nb, uint32_t* __restrict__ bins) {
    __shared__ uint32_t lbins[kMaxBins * 7];
    __shared__ uint32_t hom_node;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t s0 = blockIdx.x * blockDim.x;
    uint32_t s1 = min(s0 + blockDim.x, n) - 1;
    if (threadIdx.x == 0) {
        uint32_t a, b;
        bool va = slot_node(prims, s0, a), vb = slot_node(prims, s1, b);
        hom_node = (va && vb && a == b && a >= lv.lb && a < lv.lb + lv.K && ncount[a] >= thresh) ? a : kDone;
    }
    for (int k = threadIdx.x; k < nb * 7; k += blockDim.x) {
        int f = k % 7;
        lbins[k] = f == 0 ? 0u : (f < 4 ? fenc(kFltMax) : fenc(-kFltMax));
    }
    __syncthreads();
    uint32_t hn = hom_node;
    uint32_t node;
    if (i < n && slot_node(prims, i, node) && node >= lv.lb && node < lv.lb + lv.K && ncount[node] >= thresh) {
        float4 a = prims[2 * i], b = prims[2 * i + 1];
        int axis;
        int bk = bucket_of(a, b, cb + (node - lv.lb) * 6, nb, axis);
        uint32_t* dst = (hn != kDone) ? (lbins + bk * 7) : (bins + ((node - lv.lb) * nb + bk) * 7);
        atomicAdd(dst, 1u);
        atomicMin(dst + 1, fenc(a.x)); atomicMin(dst + 2, fenc(a.y)); atomicMin(dst + 3, fenc(a.z));
        atomicMax(dst + 4, fenc(b.x)); atomicMax(dst + 5, fenc(b.y)); atomicMax(dst + 6, fenc(b.z));
    }
    if (hn != kDone) {
        __syncthreads();
        for (int k = threadIdx.x; k < nb * 7; k += blockDim.x) {
            int bk = k / 7, f = k % 7;
            if (lbins[bk * 7] == 0) continue;
            uint32_t* dst = bins + ((hn - lv.lb) * nb + bk) * 7 + f;
            if (f == 0) atomicAdd(dst, lbins[k]);
            else if (f < 4) atomicMin(dst, lbins[k]);
            else atomicMax(dst, lbins[k]);
        }
    }
}

// One thread per node of the level: split plane sweep and decision.
__global__ void k_split(Level lv, const float4* __restrict__ nlo, const float4* __restrict__ nhi,
                        const uint32_t* __restrict__ ncount, const uint32_t* __restrict__ ndepth,
                        const uint32_t* __restrict__ cb, const uint32_t* __restrict__ bins, int nb, uint32_t thresh,
                        uint32_t max_leaf, uint32_t max_depth, SplitInfo* __restrict__ info,
                        uint32_t* __restrict__ split_flag, float4* __restrict__ child_bounds,
                        uint32_t* __restrict__ n_median) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= lv.K) return;
    uint32_t node = lv.lb + k;
    uint32_t cnt = ncount[node];
    SplitInfo si{0, 0, 0, 0};
    split_flag[k] = 0;
    if (cnt < thresh || ndepth[node] >= max_depth) {
        info[k] = si;
        return;
    }
    const uint32_t* c6 = cb + k * 6;
    float clo[3] = {fdec(c6[0]), fdec(c6[1]), fdec(c6[2])};
    float chi[3] = {fdec(c6[3]), fdec(c6[4]), fdec(c6[5])};
    int dim = max_dim(clo, chi);
    const uint32_t* bb = bins + static_cast<size_t>(k) * nb * 7;
    const int ns = nb - 1;
    float costs[kMaxBins];
    // accel/bvh_builder.cpp:146-162
    uint32_t cfwd = 0;
    float flo[3] = {kFltMax, kFltMax, kFltMax}, fhi[3] = {-kFltMax, -kFltMax, -kFltMax};
    for (int i = 0; i < ns; ++i) {
        const uint32_t* e = bb + i * 7;
        cfwd += e[0];
        for (int q = 0; q < 3; ++q) { flo[q] = fminf(flo[q], fdec(e[1 + q])); fhi[q] = fmaxf(fhi[q], fdec(e[4 + q])); }
        costs[i] = 0.0f + static_cast<float>(cfwd) * sa6(flo, fhi);
    }
    uint32_t cbwd = 0;
    float blo[3] = {kFltMax, kFltMax, kFltMax}, bhi[3] = {-kFltMax, -kFltMax, -kFltMax};
    for (int i = ns - 1; i >= 0; --i) {
        const uint32_t* e = bb + (i + 1) * 7;
        cbwd += e[0];
        for (int q = 0; q < 3; ++q) { blo[q] = fminf(blo[q], fdec(e[1 + q])); bhi[q] = fmaxf(bhi[q], fdec(e[4 + q])); }
        costs[i] = costs[i] + static_cast<float>(cbwd) * sa6(blo, bhi);
    }
    int mb = 0;
    float mc = costs[0];
    for (int i = 1; i < ns; ++i)
        if (costs[i] < mc) { mc = costs[i]; mb = i; }
    float4 l4 = nlo[node], h4 = nhi[node];
    float llo[3] = {l4.x, l4.y, l4.z}, lhi[3] = {h4.x, h4.y, h4.z};
    float lsa = sa6(llo, lhi);
    float leaf_cost = static_cast<float>(cnt);
    float min_cost = 0.5f + (lsa > 0.0f ? mc / lsa : 0.0f);
    bool capped = max_leaf > 0 && cnt > max_leaf;
    if (leaf_cost < min_cost && !capped) {
        info[k] = si;
        return;
    }
    // left/right counts and bounds from the buckets
    uint32_t lc = 0;
    float alo[3] = {kFltMax, kFltMax, kFltMax}, ahi[3] = {-kFltMax, -kFltMax, -kFltMax};
    float rlo[3] = {kFltMax, kFltMax, kFltMax}, rhi[3] = {-kFltMax, -kFltMax, -kFltMax};
    for (int i = 0; i < nb; ++i) {
        const uint32_t* e = bb + i * 7;
        if (e[0] == 0) continue;
        float* tlo = i <= mb ? alo : rlo;
        float* thi = i <= mb ? ahi : rhi;
        if (i <= mb) lc += e[0];
        for (int q = 0; q < 3; ++q) { tlo[q] = fminf(tlo[q], fdec(e[1 + q])); thi[q] = fmaxf(thi[q], fdec(e[4 + q])); }
    }
    if (lc == cnt || lc == 0) {
        if (!capped) {  // bvh_builder.cpp:104-108: partition put everything left -> leaf
            info[k] = si;
            return;
        }
        si.mode = 2;  // cannot split by buckets but the leaf cap forbids a leaf: median by order
        si.left_count = cnt / 2;
        atomicAdd(n_median, 1u);
        // child bounds are recomputed from the primitives by the caller (flagged by w = -1)
        child_bounds[4 * k + 0] = make_float4(0, 0, 0, -1.0f);
    } else {
        si.mode = 1;
        si.axis = dim;
        si.bucket = mb;
        si.left_count = lc;
        child_bounds[4 * k + 0] = make_float4(alo[0], alo[1], alo[2], 0.0f);
        child_bounds[4 * k + 1] = make_float4(ahi[0], ahi[1], ahi[2], 0.0f);
        child_bounds[4 * k + 2] = make_float4(rlo[0], rlo[1], rlo[2], 0.0f);
        child_bounds[4 * k + 3] = make_float4(rhi[0], rhi[1], rhi[2], 0.0f);
    }
    info[k] = si;
    split_flag[k] = 1;
}

__global__ void k_make_children(Level lv, uint32_t child_base, const SplitInfo* __restrict__ info,
                                const uint32_t* __restrict__ rank, const float4* __restrict__ child_bounds,
                                float4* __restrict__ nlo, float4* __restrict__ nhi, uint32_t* __restrict__ nstart,
                                uint32_t* __restrict__ ncount, uint32_t* __restrict__ nleft,
                                uint32_t* __restrict__ ndepth, uint32_t* __restrict__ stats) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= lv.K) return;
    uint32_t node = lv.lb + k;
    SplitInfo si = info[k];
    if (si.mode == 0) {
        nleft[node] = 0;
        atomicMax(&stats[0], ndepth[node]);
        atomicMax(&stats[1], ncount[node]);
        return;
    }
    uint32_t c = child_base + 2 * rank[k];
    nleft[node] = c;
    uint32_t st = nstart[node], cnt = ncount[node], d = ndepth[node] + 1;
    nstart[c] = st;
    ncount[c] = si.left_count;
    nstart[c + 1] = st + si.left_count;
    ncount[c + 1] = cnt - si.left_count;
    ndepth[c] = d;
    ndepth[c + 1] = d;
    if (si.mode == 1) {
        nlo[c] = child_bounds[4 * k + 0];
        nhi[c] = child_bounds[4 * k + 1];
        nlo[c + 1] = child_bounds[4 * k + 2];
        nhi[c + 1] = child_bounds[4 * k + 3];
    } else {
        // filled by k_node_bounds over the children (mode 0) in the next pass
        nlo[c] = make_float4(kFltMax, kFltMax, kFltMax, -1.0f);
        nlo[c + 1] = make_float4(kFltMax, kFltMax, kFltMax, -1.0f);
    }
}

// Left flag of every slot of a split node.
__global__ void k_flags(const float4* __restrict__ prims, uint32_t n, Level lv, const uint32_t* __restrict__ nstart,
                        const SplitInfo* __restrict__ info, const uint32_t* __restrict__ cb, int nb,
                        uint32_t* __restrict__ flags) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t node, f = 0;
    if (slot_node(prims, i, node) && node >= lv.lb && node < lv.lb + lv.K) {
        uint32_t k = node - lv.lb;
        SplitInfo si = info[k];
        if (si.mode == 1) {
            int axis;
            int bk = bucket_of(prims[2 * i], prims[2 * i + 1], cb + k * 6, nb, axis);
            f = bk <= si.bucket ? 1u : 0u;
        } else if (si.mode == 2) {
            f = (i - nstart[node]) < si.left_count ? 1u : 0u;
        }
    }
    flags[i] = f;
}

__global__ void k_scatter(const float4* __restrict__ src, float4* __restrict__ dst, uint32_t n, Level lv,
                          const uint32_t* __restrict__ nstart, const uint32_t* __restrict__ nleft,
                          const SplitInfo* __restrict__ info, const uint32_t* __restrict__ flags,
                          const uint32_t* __restrict__ scan) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 a = src[2 * i], b = src[2 * i + 1];
    uint32_t node = __float_as_uint(b.w);
    uint32_t pos = i;
    if (node != kDone && node >= lv.lb && node < lv.lb + lv.K) {
        SplitInfo si = info[node - lv.lb];
        if (si.mode == 0) {
            b.w = __uint_as_float(kDone);
        } else {
            uint32_t st = nstart[node];
            uint32_t lrank = scan[i] - scan[st];
            uint32_t c = nleft[node];
            if (flags[i]) {
                pos = st + lrank;
                b.w = __uint_as_float(c);
            } else {
                pos = st + si.left_count + ((i - st) - lrank);
                b.w = __uint_as_float(c + 1);
            }
        }
    }
    dst[2 * pos] = a;
    dst[2 * pos + 1] = b;
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
    size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Init the per-level bound slots: min fields to enc(+FLT_MAX), max to enc(-FLT_MAX).
__global__ void k_init_bounds6(uint32_t* p, uint32_t count6, uint32_t enc_max, uint32_t enc_min) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count6 * 6) p[i] = (i % 6) < 3 ? enc_max : enc_min;
}
__global__ void k_init_bins(uint32_t* p, size_t nwords, uint32_t enc_max, uint32_t enc_min) {
    size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < nwords) {
        int f = static_cast<int>(i % 7);
        p[i] = f == 0 ? 0u : (f < 4 ? enc_max : enc_min);
    }
}

// Nodes whose bounds were flagged for recomputation (median splits, root)
// take the prim-bounds reduction result.
__global__ void k_take_bounds(Level lv, const uint32_t* __restrict__ b6, float4* __restrict__ nlo,
                              float4* __restrict__ nhi, int force) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= lv.K) return;
    uint32_t node = lv.lb + k;
    if (!force && nlo[node].w != -1.0f) return;
    const uint32_t* e = b6 + k * 6;
    nlo[node] = make_float4(fdec(e[0]), fdec(e[1]), fdec(e[2]), 0.0f);
    nhi[node] = make_float4(fdec(e[3]), fdec(e[4]), fdec(e[5]), 0.0f);
}

__global__ void k_emit_nodes(uint32_t m, const float4* __restrict__ nlo, const float4* __restrict__ nhi,
                             const uint32_t* __restrict__ nstart, const uint32_t* __restrict__ ncount,
                             const uint32_t* __restrict__ nleft, float4* __restrict__ out) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    if (k == 1) {  // unused slot
        out[2] = make_float4(0, 0, 0, 0);
        out[3] = make_float4(0, 0, 0, 0);
        return;
    }
    float4 lo = nlo[k], hi = nhi[k];
    uint32_t l = nleft[k];
    if (l != 0) {
        lo.w = __uint_as_float(l);
        hi.w = __uint_as_float(l + 1);
    } else {
        lo.w = __uint_as_float(kLeafFlag | nstart[k]);
        hi.w = __uint_as_float(ncount[k]);
    }
    out[2 * k] = lo;
    out[2 * k + 1] = hi;
}

__global__ void k_prim_order(const float4* __restrict__ prims, uint32_t n, uint32_t* __restrict__ order) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) order[i] = __float_as_uint(prims[2 * i].w);
}

// ---- device-wide exclusive scan (2048 items per block, two levels) ----

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds_waves, uint32_t* total) {
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = v;
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) lds_waves[w] = x;
    __syncthreads();
    uint32_t wave_off = 0, tot = 0;
    for (int q = 0; q < kBlock / 64; ++q) {
        uint32_t s = lds_waves[q];
        if (q < w) wave_off += s;
        tot += s;
    }
    *total = tot;
    return wave_off + x - v;
}

__global__ void k_scan_blocks(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint32_t n,
                              uint32_t* __restrict__ block_sums) {
    __shared__ uint32_t waves[kBlock / 64];
    uint32_t base = blockIdx.x * kScanBlock + threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
    for (int j = 0; j < kScanItems; ++j) {
        uint32_t idx = base + j;
        v[j] = idx < n ? in[idx] : 0u;
        s += v[j];
    }
    uint32_t tot;
    uint32_t ex = block_exclusive_scan(s, waves, &tot);
    for (int j = 0; j < kScanItems; ++j) {
        uint32_t idx = base + j;
        if (idx < n) out[idx] = ex;
        ex += v[j];
    }
    if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

__global__ void k_scan_sums(uint32_t* __restrict__ sums, uint32_t nb, uint32_t* __restrict__ total) {
    __shared__ uint32_t waves[kBlock / 64];
    uint32_t base = threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
    for (int j = 0; j < kScanItems; ++j) {
        v[j] = (base + j) < nb ? sums[base + j] : 0u;
        s += v[j];
    }
    uint32_t tot;
    uint32_t ex = block_exclusive_scan(s, waves, &tot);
    __syncthreads();
    for (int j = 0; j < kScanItems; ++j) {
        if (base + j < nb) sums[base + j] = ex;
        ex += v[j];
    }
    if (threadIdx.x == 0 && total) *total = tot;
}

__global__ void k_scan_add(uint32_t* __restrict__ out, uint32_t n, const uint32_t* __restrict__ sums) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += sums[i / kScanBlock];
}

inline uint32_t nblocks(size_t n, int b = kBlock) { return static_cast<uint32_t>((n + b - 1) / b); }

#define HIP_OK(expr)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) {                                                              \
            if (err) *err = std::string(#expr) + ": " + hipGetErrorString(e_);               \
            ok = false;                                                                      \
            goto cleanup;                                                                    \
        }                                                                                    \
    } while (0)

}  // namespace

void scan_exclusive_u32(const uint32_t* d_in, uint32_t* d_out, uint32_t n, uint32_t* d_tmp, uint32_t* d_total,
                        hipStream_t stream) {
    if (n == 0) {
        if (d_total) hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(1), 0, stream, d_total, 0u, size_t(1));
        return;
    }
    uint32_t nb = nblocks(n, kScanBlock);
    hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(kBlock), 0, stream, d_in, d_out, n, d_tmp);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kBlock), 0, stream, d_tmp, nb, d_total);
    hipLaunchKernelGGL(k_scan_add, dim3(nblocks(n)), dim3(kBlock), 0, stream, d_out, n, d_tmp);
}

bool gpu_build_bvh(const float4* d_vpos, const uint4* d_faces, uint32_t n, const BvhBuildOptions& opt,
                   hipStream_t stream, BvhDeviceResult* res, BvhHostTopology* topo, std::string* err) {
    bool ok = true;
    *res = BvhDeviceResult();
    const int nb = opt.bins;
    const uint32_t thresh = static_cast<uint32_t>(std::max(1, opt.leaf_threshold));
    const uint32_t max_leaf = static_cast<uint32_t>(std::max(0, opt.max_leaf));
    const uint32_t max_depth = kStackSize;
    const uint32_t cap = 2 * n + 2;
    const uint32_t maxK = n + 1;  // nodes per level never exceed the primitive count
    const uint32_t enc_max = fenc_host(kFltMax), enc_min = fenc_host(-kFltMax);

    float4 *prims_a = nullptr, *prims_b = nullptr, *nlo = nullptr, *nhi = nullptr, *child_bounds = nullptr;
    uint32_t *nstart = nullptr, *ncount = nullptr, *nleft = nullptr, *ndepth = nullptr;
    uint32_t *cb = nullptr, *bins = nullptr, *split_flag = nullptr, *rank = nullptr, *pflag = nullptr,
             *pscan = nullptr, *tmp = nullptr, *scal = nullptr;
    SplitInfo* info = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    uint32_t level_begin = 0, level_count = 1, next_free = 2;
    uint32_t depth_stats[2] = {0, 0};
    bool need_prim_bounds = false;

    if (nb < 2 || nb > kMaxBins) {
        if (err) *err = "bvh bins must be in [2, 32]";
        return false;
    }
    if (n > 4u * 1024u * 1024u) {
        if (err) *err = "GPU BVH build supports at most 4,194,304 triangles";
        return false;
    }
    if (n == 0) return true;

    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    HIP_OK(hipEventRecord(e0));
    HIP_OK(hipMalloc(reinterpret_cast<void**>(&prims_a), sizeof(float4) * 2 * n));
    HIP_OK(hipMalloc(reinterpret_cast<void**>(&prims_b), sizeof(float4) * 2 * n));
    HIP_OK(hipMalloc(reinterpret_cast<void**>(&nlo), sizeof(float4) * cap));
    HIP_OK(hipMalloc(reinterpret_cast<void**>(&nhi), sizeof(float4) * cap));
    HIP_OK(hipMalloc(reinterpret_cast<void**>(&nstart), sizeof(uint32_t) * cap));
    HIP_OK(hipMalloc(reinterpret_cast<void**>(&ncount), sizeof(uint32_t) * cap));
    HIP_OK(hipMalloc(reinterpret_cast<void**>(&nleft), sizeof(uint32_t) * cap));
    HIP_OK(hipMalloc(reinterpret_cast<void**>(&ndepth), sizeof(uint32_t) * cap));
    HIP_OK(hipMalloc(reinterpret_cast<void**>(&cb), sizeof(uint32_t) * 6 * maxK));
    HIP_OK(hipMalloc(reinterpret_cast<void**>(&bins), sizeof(uint32_t) * 7 * nb * static_cast<size_t>(maxK)));
    HIP_OK(hipMalloc(reinterpret_cast<void**>(&split_flag), sizeof(uint32_t) * maxK));
    HIP_OK(hipMalloc(reinterpret_cast<void**>(&rank), sizeof(uint32_t) * maxK));
    HIP_OK(hipMalloc(reinterpret_cast<void**>(&info), sizeof(SplitInfo) * maxK));
    HIP_OK(hipMalloc(reinterpret_cast<void**>(&child_bounds), sizeof(float4) * 4 * maxK));
    HIP_OK(hipMalloc(reinterpret_cast<void**>(&pflag), sizeof(uint32_t) * n));
    HIP_OK(hipMalloc(reinterpret_cast<void**>(&pscan), sizeof(uint32_t) * n));
    HIP_OK(hipMalloc(reinterpret_cast<void**>(&tmp), sizeof(uint32_t) * (kScanBlock + 32)));
    HIP_OK(hipMalloc(reinterpret_cast<void**>(&scal), sizeof(uint32_t) * 4));
    HIP_OK(hipMemsetAsync(scal, 0, sizeof(uint32_t) * 4, stream));

    hipLaunchKernelGGL(k_prim_init, dim3(nblocks(n)), dim3(kBlock), 0, stream, d_vpos, d_faces, n, prims_a);
    {   // root record: [0, n), depth 0, bounds = union of primitive bounds
        uint32_t zero = 0;
        HIP_OK(hipMemcpyAsync(nstart, &zero, 4, hipMemcpyHostToDevice, stream));
        HIP_OK(hipMemcpyAsync(ncount, &n, 4, hipMemcpyHostToDevice, stream));
        HIP_OK(hipMemcpyAsync(ndepth, &zero, 4, hipMemcpyHostToDevice, stream));
        Level root{0, 1};
        hipLaunchKernelGGL(k_init_bounds6, dim3(1), dim3(kBlock), 0, stream, cb, 1u, enc_max, enc_min);
        hipLaunchKernelGGL(k_node_bounds, dim3(nblocks(n)), dim3(kBlock), 0, stream, prims_a, n, root, 0, ncount,
                           thresh, cb);
        hipLaunchKernelGGL(k_take_bounds, dim3(1), dim3(kBlock), 0, stream, root, cb, nlo, nhi, 1);
    }
    while (level_count > 0) {
        res->level_begin.push_back(level_begin);
        res->level_count.push_back(level_count);
        Level lv{level_begin, level_count};
        const uint32_t K = level_count;
        if (need_prim_bounds) {  // median-split children take their bounds from the primitives
            hipLaunchKernelGGL(k_init_bounds6, dim3(nblocks(static_cast<size_t>(K) * 6)), dim3(kBlock), 0, stream, cb,
                               K, enc_max, enc_min);
            hipLaunchKernelGGL(k_node_bounds, dim3(nblocks(n)), dim3(kBlock), 0, stream, prims_a, n, lv, 0, ncount,
                               thresh, cb);
            hipLaunchKernelGGL(k_take_bounds, dim3(nblocks(K)), dim3(kBlock), 0, stream, lv, cb, nlo, nhi, 0);
        }
        // centroid bounds
        hipLaunchKernelGGL(k_init_bounds6, dim3(nblocks(static_cast<size_t>(K) * 6)), dim3(kBlock), 0, stream, cb, K,
                           enc_max, enc_min);
        hipLaunchKernelGGL(k_node_bounds, dim3(nblocks(n)), dim3(kBlock), 0, stream, prims_a, n, lv, 1, ncount,
                           thresh, cb);
        size_t bin_words = static_cast<size_t>(K) * nb * 7;
        hipLaunchKernelGGL(k_init_bins, dim3(nblocks(bin_words)), dim3(kBlock), 0, stream, bins, bin_words, enc_max,
                           enc_min);
        hipLaunchKernelGGL(k_bin, dim3(nblocks(n)), dim3(kBlock), 0, stream, prims_a, n, lv, ncount, thresh, cb, nb,
                           bins);
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(1), 0, stream, scal + 1, 0u, size_t(1));
        hipLaunchKernelGGL(k_split, dim3(nblocks(K)), dim3(kBlock), 0, stream, lv, nlo, nhi, ncount, ndepth, cb, bins,
                           nb, thresh, max_leaf, max_depth, info, split_flag, child_bounds, scal + 1);
        scan_exclusive_u32(split_flag, rank, K, tmp, scal, stream);
        uint32_t counts2[2] = {0, 0};
        HIP_OK(hipMemcpyAsync(counts2, scal, 8, hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        const uint32_t nsplit = counts2[0];
        need_prim_bounds = counts2[1] > 0;
        hipLaunchKernelGGL(k_make_children, dim3(nblocks(K)), dim3(kBlock), 0, stream, lv, next_free, info, rank,
                           child_bounds, nlo, nhi, nstart, ncount, nleft, ndepth, scal + 2);
        hipLaunchKernelGGL(k_flags, dim3(nblocks(n)), dim3(kBlock), 0, stream, prims_a, n, lv, nstart, info, cb, nb,
                           pflag);
        scan_exclusive_u32(pflag, pscan, n, tmp, nullptr, stream);
        hipLaunchKernelGGL(k_scatter, dim3(nblocks(n)), dim3(kBlock), 0, stream, prims_a, prims_b, n, lv, nstart,
                           nleft, info, pflag, pscan);
        std::swap(prims_a, prims_b);
        level_begin = next_free;
        level_count = 2 * nsplit;
        next_free += 2 * nsplit;
        if (next_free > cap) {
            if (err) *err = "BVH node capacity exceeded";
            ok = false;
            goto cleanup;
        }
    }
    HIP_OK(hipGetLastError());
    res->num_nodes = next_free;
    HIP_OK(hipMalloc(reinterpret_cast<void**>(&res->nodes), sizeof(float4) * 2 * res->num_nodes));
    HIP_OK(hipMalloc(reinterpret_cast<void**>(&res->prim_order), sizeof(uint32_t) * n));
    hipLaunchKernelGGL(k_emit_nodes, dim3(nblocks(res->num_nodes)), dim3(kBlock), 0, stream, res->num_nodes, nlo, nhi,
                       nstart, ncount, nleft, res->nodes);
    hipLaunchKernelGGL(k_prim_order, dim3(nblocks(n)), dim3(kBlock), 0, stream, prims_a, n, res->prim_order);
    HIP_OK(hipMemcpyAsync(depth_stats, scal + 2, 8, hipMemcpyDeviceToHost, stream));
    HIP_OK(hipEventRecord(e1, stream));
    HIP_OK(hipStreamSynchronize(stream));
    HIP_OK(hipGetLastError());
    res->depth = depth_stats[0];
    res->max_leaf = depth_stats[1];
    HIP_OK(hipEventElapsedTime(&res->build_ms, e0, e1));
    if (topo) {
        const uint32_t m = res->num_nodes;
        std::vector<float4> lo(m), hi(m);
        topo->num_nodes = m;
        topo->depth = res->depth;
        topo->max_leaf = res->max_leaf;
        topo->left.resize(m);
        topo->start.resize(m);
        topo->count.resize(m);
        topo->prim_order.resize(n);
        HIP_OK(hipMemcpy(lo.data(), nlo, sizeof(float4) * m, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(hi.data(), nhi, sizeof(float4) * m, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(topo->left.data(), nleft, 4 * m, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(topo->start.data(), nstart, 4 * m, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(topo->count.data(), ncount, 4 * m, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(topo->prim_order.data(), res->prim_order, 4 * n, hipMemcpyDeviceToHost));
        topo->lo.resize(3 * m);
        topo->hi.resize(3 * m);
        for (uint32_t k = 0; k < m; ++k) {
            topo->lo[3 * k] = lo[k].x; topo->lo[3 * k + 1] = lo[k].y; topo->lo[3 * k + 2] = lo[k].z;
            topo->hi[3 * k] = hi[k].x; topo->hi[3 * k + 1] = hi[k].y; topo->hi[3 * k + 2] = hi[k].z;
        }
        if (m > 1) topo->left[1] = 0, topo->count[1] = 0, topo->start[1] = 0;
    }
cleanup:
    (void)hipStreamSynchronize(stream);
    for (void* p : {static_cast<void*>(prims_a), static_cast<void*>(prims_b), static_cast<void*>(nlo),
                    static_cast<void*>(nhi), static_cast<void*>(nstart), static_cast<void*>(ncount),
                    static_cast<void*>(nleft), static_cast<void*>(ndepth), static_cast<void*>(cb),
                    static_cast<void*>(bins), static_cast<void*>(split_flag), static_cast<void*>(rank),
                    static_cast<void*>(info), static_cast<void*>(child_bounds), static_cast<void*>(pflag),
                    static_cast<void*>(pscan), static_cast<void*>(tmp), static_cast<void*>(scal)})
        if (p) (void)hipFree(p);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (!ok) {
        if (res->nodes) (void)hipFree(res->nodes);
        if (res->prim_order) (void)hipFree(res->prim_order);
        *res = BvhDeviceResult();
    }
    return ok;
}

}  // namespace ptd

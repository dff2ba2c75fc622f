// BVH traversal over the sibling-pair node layout (see dev_scene.h).
//
// Both children of an internal node are fetched with one 64 B read and
// slab-tested against the current closest distance; the nearer hit child is
// entered and the farther one deferred on a per-thread stack held in LDS
// (stack[level * stride + lane], conflict-free).  Leaves are ranges of
// triangles ended by the "last" bit.  The slab test and the triangle test use
// the reference's arithmetic (Part_2/src/accel/bvh_traversal.cuh:29-41,
// geometry/triangle.h:13-76): the result is the same closest triangle (strict
// '<' keeps the first found among exact ties, so only exact ties can resolve
// differently from the reference's order).
#pragma once

#include "dev_math.h"
#include "dev_scene.h"

namespace ptd {

struct Hit {
    float t;
    float u, v;
    uint32_t tri;  // BVH slot; 0xFFFFFFFF on a miss
};

struct TraceCounters {
    uint32_t nodes, tris;
};

struct RayInv {
    F3 o, inv;
};

__device__ __forceinline__ RayInv make_ray_inv(F3 o, F3 d) {
    RayInv r;
    r.o = o;
    r.inv = f3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    return r;
}

// bvh_traversal.cuh:29-41 (t_min starts at 0, t_max at the current distance);
// returns the entry distance, or a negative value on a miss.
__device__ __forceinline__ float slab(const RayInv& r, float4 lo, float4 hi, float tmax) {
    float t_min = 0.0f, t_max = tmax;
    float t1 = (lo.x - r.o.x) * r.inv.x, t2 = (hi.x - r.o.x) * r.inv.x;
    t_min = fminf(fmaxf(t1, t_min), fmaxf(t2, t_min));
    t_max = fmaxf(fminf(t1, t_max), fminf(t2, t_max));
    t1 = (lo.y - r.o.y) * r.inv.y; t2 = (hi.y - r.o.y) * r.inv.y;
    t_min = fminf(fmaxf(t1, t_min), fmaxf(t2, t_min));
    t_max = fmaxf(fminf(t1, t_max), fminf(t2, t_max));
    t1 = (lo.z - r.o.z) * r.inv.z; t2 = (hi.z - r.o.z) * r.inv.z;
    t_min = fminf(fmaxf(t1, t_min), fmaxf(t2, t_min));
    t_max = fmaxf(fminf(t1, t_max), fminf(t2, t_max));
    return t_min <= t_max ? t_min : -1.0f;
}

// Moller-Trumbore (triangle.h:13-51 / :53-76).  Returns t, or -1 on a miss;
// u, v are the barycentrics of p1 and p2.
__device__ __forceinline__ float tri_test(F3 o, F3 d, const float4* tri, float& u, float& v, uint32_t& word) {
    // no FMA contraction: every product and sum rounds as in the reference's
    // float arithmetic, so t matches it to the last bit
#pragma clang fp contract(off)
    float4 a = tri[0], b = tri[1], c = tri[2];
    word = __float_as_uint(a.w);
    F3 e1 = f3(b.x, b.y, b.z), e2 = f3(c.x, c.y, c.z);
    F3 pvec = cross_rn(d, e2);
    float det = dot_rn(e1, pvec);
    if (fabsf(det) < kEps) return -1.0f;
    float inv_det = 1.0f / det;
    F3 tvec = o - f3(a.x, a.y, a.z);
    u = dot_rn(tvec, pvec) * inv_det;
    if (u < 0.0f || u > 1.0f) return -1.0f;
    F3 qvec = cross_rn(tvec, e1);
    v = dot_rn(d, qvec) * inv_det;
    if (v < 0.0f || u + v > 1.0f) return -1.0f;
    return dot_rn(e2, qvec) * inv_det;
}

// Per-thread traversal stack: the first LDS entries live in LDS
// (stk[level * stride]), deeper ones in global memory
// (spill[(level - LDS) * spill_stride]); kStackSize entries in all.  With
// LDS == kStackSize there is no spill area.
template <int LDS>
struct TraceStack {
    uint32_t* stk;
    int stride;
    uint32_t* spill;
    uint32_t spill_stride;
    __device__ __forceinline__ void put(int level, uint32_t v) const {
        if (LDS >= kStackSize || level < LDS) stk[level * stride] = v;
        else spill[static_cast<size_t>(level - LDS) * spill_stride] = v;
    }
    __device__ __forceinline__ uint32_t get(int level) const {
        if (LDS >= kStackSize || level < LDS) return stk[level * stride];
        return spill[static_cast<size_t>(level - LDS) * spill_stride];
    }
};

// Closest hit.
template <bool COUNT, int LDS>
__device__ __forceinline__ Hit trace_closest_s(const DevScene& s, F3 o, F3 d, const TraceStack<LDS>& sk,
                                               TraceCounters& cnt, uint32_t& overflow) {
    Hit h;
    h.t = kInf;
    h.u = h.v = 0.0f;
    h.tri = 0xFFFFFFFFu;
    if (s.num_nodes == 0) return h;
    RayInv r = make_ray_inv(o, d);
    float4 rlo = s.nodes[0], rhi = s.nodes[1];
    if (COUNT) cnt.nodes += 1;
    if (slab(r, rlo, rhi, h.t) < 0.0f) return h;
    uint32_t cur = __float_as_uint(rlo.w);
    int top = 0;
    while (true) {
        if (cur & kLeafFlag) {
            uint32_t i = cur & ~kLeafFlag;
            while (true) {
                float u, v;
                uint32_t word;
                float t = tri_test(o, d, s.tris + 3 * i, u, v, word);
                if (COUNT) cnt.tris += 1;
                // triangle.h:38 rejects t < 1e-4; bvh_traversal.cuh:72 keeps strict improvements
                if (t >= 1e-4f && t < h.t) {
                    h.t = t;
                    h.u = u;
                    h.v = v;
                    h.tri = i;
                }
                if (word & 0x80000000u) break;
                ++i;
            }
        } else {
            const float4* c = s.nodes + 2 * cur;
            float4 alo = c[0], ahi = c[1], blo = c[2], bhi = c[3];
            if (COUNT) cnt.nodes += 2;
            float ta = slab(r, alo, ahi, h.t);
            float tb = slab(r, blo, bhi, h.t);
            uint32_t da = __float_as_uint(alo.w), db = __float_as_uint(blo.w);
            if (ta >= 0.0f && tb >= 0.0f) {
                uint32_t near = ta <= tb ? da : db, far = ta <= tb ? db : da;
                if (top < kStackSize) sk.put(top++, far);
                else overflow++;
                cur = near;
                continue;
            }
            if (ta >= 0.0f) { cur = da; continue; }
            if (tb >= 0.0f) { cur = db; continue; }
        }
        if (top == 0) break;
        cur = sk.get(--top);
    }
    return h;
}

// Any hit in (1e-4, tmax) (occludedScene, bvh_traversal.cuh:98-133).
template <bool COUNT, int LDS>
__device__ __forceinline__ bool trace_any_s(const DevScene& s, F3 o, F3 d, float tmax, const TraceStack<LDS>& sk,
                                            TraceCounters& cnt, uint32_t& overflow) {
    if (s.num_nodes == 0) return false;
    RayInv r = make_ray_inv(o, d);
    float4 rlo = s.nodes[0], rhi = s.nodes[1];
    if (COUNT) cnt.nodes += 1;
    if (slab(r, rlo, rhi, tmax) < 0.0f) return false;
    uint32_t cur = __float_as_uint(rlo.w);
    int top = 0;
    while (true) {
        if (cur & kLeafFlag) {
            uint32_t i = cur & ~kLeafFlag;
            while (true) {
                float u, v;
                uint32_t word;
                float t = tri_test(o, d, s.tris + 3 * i, u, v, word);
                if (COUNT) cnt.tris += 1;
                if (t > 1e-4f && t < tmax) return true;
                if (word & 0x80000000u) break;
                ++i;
            }
        } else {
            const float4* c = s.nodes + 2 * cur;
            float4 alo = c[0], ahi = c[1], blo = c[2], bhi = c[3];
            if (COUNT) cnt.nodes += 2;
            float ta = slab(r, alo, ahi, tmax);
            float tb = slab(r, blo, bhi, tmax);
            uint32_t da = __float_as_uint(alo.w), db = __float_as_uint(blo.w);
            if (ta >= 0.0f && tb >= 0.0f) {
                uint32_t near = ta <= tb ? da : db, far = ta <= tb ? db : da;
                if (top < kStackSize) sk.put(top++, far);
                else overflow++;
                cur = near;
                continue;
            }
            if (ta >= 0.0f) { cur = da; continue; }
            if (tb >= 0.0f) { cur = db; continue; }
        }
        if (top == 0) break;
        cur = sk.get(--top);
    }
    return false;
}

// The all-LDS forms (stk: this thread's first LDS slot, kStackSize entries).
template <bool COUNT>
__device__ __forceinline__ Hit trace_closest(const DevScene& s, F3 o, F3 d, uint32_t* stk, int stride,
                                             TraceCounters& cnt, uint32_t& overflow) {
    TraceStack<kStackSize> sk{stk, stride, nullptr, 0};
    return trace_closest_s<COUNT>(s, o, d, sk, cnt, overflow);
}
template <bool COUNT>
__device__ __forceinline__ bool trace_any(const DevScene& s, F3 o, F3 d, float tmax, uint32_t* stk, int stride,
                                          TraceCounters& cnt, uint32_t& overflow) {
    TraceStack<kStackSize> sk{stk, stride, nullptr, 0};
    return trace_any_s<COUNT>(s, o, d, tmax, sk, cnt, overflow);
}

}  // namespace ptd

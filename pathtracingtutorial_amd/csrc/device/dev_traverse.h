// BVH traversal over the sibling-pair node layout (see dev_scene.h).
//
// Both children of an internal node are fetched with one 64 B read and
// slab-tested against the current closest distance; the nearer hit child is
// entered and the farther one deferred on a per-thread stack held in LDS
// (stack[level * stride + lane], conflict-free).  Leaves are ranges of
// triangles ended by the "last" bit.  The slab test and the triangle test use
// the reference's arithmetic (Part_2/src/accel/bvh_traversal.cuh:29-41,
// geometry/triangle.h:13-76), so t is the reference's to the bit.  Children
// are entered by slab entry distance (the reference orders them by child
// centroid along the parent's widest axis, bvh_traversal.cuh:78-91), and the
// strict '<' keeps the first triangle found among equal distances: the
// closest triangle is the reference's except among exact-distance ties,
// which may resolve the other way.
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
    F3 oi;  // 2^-20 |o inv| per axis: part of the wide node test's rounding slack
};

__device__ __forceinline__ RayInv make_ray_inv(F3 o, F3 d) {
    RayInv r;
    r.o = o;
    r.inv = f3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    r.oi = f3(fabsf(o.x * r.inv.x) * 0x1p-20f, fabsf(o.y * r.inv.y) * 0x1p-20f, fabsf(o.z * r.inv.z) * 0x1p-20f);
    return r;
}

// Scene arrays live in device memory; loading them through a global-address-
// space pointer gives global loads (a generic pointer read from the job table
// compiles to flat loads, which also count against the LDS counter that the
// traversal stack's LDS accesses wait on).
// (The load goes through a plain vector type: float4's copy constructor
// takes a generic reference and would undo the address space.)
typedef float V4F __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 gld(const float4* p) {
    const V4F v = *(const __attribute__((address_space(1))) V4F*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}

__device__ __forceinline__ uint32_t gld_u32(const uint32_t* p) {
    return *(const __attribute__((address_space(1))) uint32_t*)p;
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

// The same test in fewer instructions.  When every component of inv is
// finite, no product above is NaN, and then fmin(fmax(t1, m), fmax(t2, m)) =
// max(min(t1, t2), m) and fmax(fmin(t1, M), fmin(t2, M)) = min(max(t1, t2), M)
// exactly (min and max do not round; the per-axis updates commute), so entry
// and exit are the reference's to the bit.  A zero (or subnormal) direction
// component gives inf and possibly 0 * inf = NaN, which the reference's form
// turns into "no constraint on this axis": such rays use slab().
__device__ __forceinline__ float slab_fast(const RayInv& r, float4 lo, float4 hi, float tmax) {
    const float x1 = (lo.x - r.o.x) * r.inv.x, x2 = (hi.x - r.o.x) * r.inv.x;
    const float y1 = (lo.y - r.o.y) * r.inv.y, y2 = (hi.y - r.o.y) * r.inv.y;
    const float z1 = (lo.z - r.o.z) * r.inv.z, z2 = (hi.z - r.o.z) * r.inv.z;
    const float tn = fmaxf(fmaxf(fminf(x1, x2), fminf(y1, y2)), fmaxf(fminf(z1, z2), 0.0f));
    const float tf = fminf(fminf(fmaxf(x1, x2), fmaxf(y1, y2)), fminf(fmaxf(z1, z2), tmax));
    return tn <= tf ? tn : -1.0f;
}

template <bool FAST>
__device__ __forceinline__ float slab_v(const RayInv& r, float4 lo, float4 hi, float tmax) {
    if constexpr (FAST) return slab_fast(r, lo, hi, tmax);
    else return slab(r, lo, hi, tmax);
}

// inv = 1 / d has no component above 2^64 (none infinite or NaN): slab_fast
// applies, and the wide test's products S inv stay finite for any node
// narrower than 2^64 per plane step.
__device__ __forceinline__ bool finite_inv(const RayInv& r) {
    return fabsf(r.inv.x) <= 0x1p64f && fabsf(r.inv.y) <= 0x1p64f && fabsf(r.inv.z) <= 0x1p64f;
}

// 1 / x correctly rounded for |x| in [2^-126, 2^126): the hardware
// reciprocal (within 1 ulp) and one Newton step with fma.  Equal to IEEE
// division for every float of either sign in that range (pt_probe_reciprocal
// compares all 2 x 252 x 2^23 of them); three instructions instead of the
// division's scale / fma / fixup sequence.
__device__ __forceinline__ float recip_rn(float x) {
    const float r0 = __builtin_amdgcn_rcpf(x);
    return fmaf(fmaf(-x, r0, 1.0f), r0, r0);
}

// Moller-Trumbore (triangle.h:13-51 / :53-76).  Returns t, or -1 on a miss;
// u, v are the barycentrics of p1 and p2.
__device__ __forceinline__ float tri_test(F3 o, F3 d, const float4* tri, float& u, float& v, uint32_t& word) {
    // no FMA contraction: every product and sum rounds as in the reference's
    // float arithmetic, so t matches it to the last bit
#pragma clang fp contract(off)
    float4 a = gld(tri), b = gld(tri + 1), c = gld(tri + 2);
    word = __float_as_uint(a.w);
    F3 e1 = f3(b.x, b.y, b.z), e2 = f3(c.x, c.y, c.z);
    F3 pvec = cross_rn(d, e2);
    float det = dot_rn(e1, pvec);
    if (fabsf(det) < kEps) return -1.0f;
    float inv_det;
    if (__builtin_expect(fabsf(det) < 0x1p126f, 1)) inv_det = recip_rn(det);
    else inv_det = 1.0f / det;
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
// (spill[(level - LDS) * spill_stride]); TOTAL entries in all.  With
// LDS == TOTAL there is no spill area.  An entry E is the node word
// (uint32_t), or for closest-hit traversals a DistEntry that also holds the
// entry distance of the node's box, so that a popped node lying beyond the
// closest hit found since it was pushed is dropped without being fetched
// (the reference re-tests a popped node against the current distance,
// bvh_traversal.cuh:58-60).
// (one 64-bit word: the node word in the low half, the distance's bits in the high half)
using DistEntry = unsigned long long;

template <int LDS, int TOTAL = kStackSize, class E = uint32_t>
struct TraceStack {
    static constexpr int kTotal = TOTAL;
    static constexpr int kLds = LDS;
    static constexpr bool kDist = sizeof(E) == sizeof(DistEntry);
    using Entry = E;
    E* stk;
    int stride;
    E* spill;
    uint32_t spill_stride;
    // (the device-memory part is addressed as global memory: a generic
    // pointer would let the compiler merge both cases into one flat access)
    __device__ __forceinline__ void put(int level, E v) const {
        if (LDS >= TOTAL || level < LDS) stk[level * stride] = v;
        else *(__attribute__((address_space(1))) E*)(spill + static_cast<size_t>(level - LDS) * spill_stride) = v;
    }
    __device__ __forceinline__ E get(int level) const {
        if (LDS >= TOTAL || level < LDS) return stk[level * stride];
        return *(const __attribute__((address_space(1))) E*)(spill + static_cast<size_t>(level - LDS) * spill_stride);
    }
    // levels top, top + 1, top + 2, all in the LDS part
    __device__ __forceinline__ void put3_lds(int top, E a, E b, E c) const {
        E* p = stk + top * stride;
        p[0] = a;
        p[stride] = b;
        p[2 * stride] = c;
    }
};

__device__ __forceinline__ uint32_t entry_word(uint32_t e) { return e; }
__device__ __forceinline__ uint32_t entry_word(DistEntry e) { return static_cast<uint32_t>(e); }
__device__ __forceinline__ float entry_dist(uint32_t) { return 0.0f; }
__device__ __forceinline__ float entry_dist(DistEntry e) { return __uint_as_float(static_cast<uint32_t>(e >> 32)); }
template <class E>
__device__ __forceinline__ E make_entry(uint32_t word, float dist) {
    if constexpr (sizeof(E) == sizeof(DistEntry))
        return (static_cast<DistEntry>(__float_as_uint(dist)) << 32) | word;
    else return word;
}

// Pops the next entry whose box can still hold a hit closer than tmax (any
// entry, for stacks without distances); false when the stack is empty.
template <class SK>
__device__ __forceinline__ bool pop_live(const SK& sk, int& top, float tmax, uint32_t& cur) {
    while (top > 0) {
        const typename SK::Entry e = sk.get(--top);
        if (!SK::kDist || entry_dist(e) <= tmax) {
            cur = entry_word(e);
            return true;
        }
    }
    return false;
}

// Closest hit.
template <bool FAST, bool COUNT, class SK>
__device__ __forceinline__ Hit trace_closest_s_impl(const DevScene& s, F3 o, F3 d, const RayInv& r,
                                                    const SK& sk, TraceCounters& cnt,
                                                    uint32_t& overflow) {
    using E = typename SK::Entry;
    constexpr int TOTAL = SK::kTotal;
    Hit h;
    h.t = kInf;
    h.u = h.v = 0.0f;
    h.tri = 0xFFFFFFFFu;
    if (s.num_nodes == 0) return h;
    float4 rlo = gld(s.nodes), rhi = gld(s.nodes + 1);
    if (COUNT) cnt.nodes += 1;
    if (slab_v<FAST>(r, rlo, rhi, h.t) < 0.0f) return h;
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
            float4 alo = gld(c), ahi = gld(c + 1), blo = gld(c + 2), bhi = gld(c + 3);
            if (COUNT) cnt.nodes += 2;
            float ta = slab_v<FAST>(r, alo, ahi, h.t);
            float tb = slab_v<FAST>(r, blo, bhi, h.t);
            uint32_t da = __float_as_uint(alo.w), db = __float_as_uint(blo.w);
            if (ta >= 0.0f && tb >= 0.0f) {
                uint32_t near = ta <= tb ? da : db, far = ta <= tb ? db : da;
                if (top < TOTAL) sk.put(top++, make_entry<E>(far, fmaxf(ta, tb)));
                else overflow++;
                cur = near;
                continue;
            }
            if (ta >= 0.0f) { cur = da; continue; }
            if (tb >= 0.0f) { cur = db; continue; }
        }
        if (!pop_live(sk, top, h.t, cur)) break;
    }
    return h;
}

template <bool COUNT, class SK>
__device__ __forceinline__ Hit trace_closest_s(const DevScene& s, F3 o, F3 d, const SK& sk,
                                               TraceCounters& cnt, uint32_t& overflow) {
    const RayInv r = make_ray_inv(o, d);
    if (finite_inv(r)) return trace_closest_s_impl<true, COUNT>(s, o, d, r, sk, cnt, overflow);
    return trace_closest_s_impl<false, COUNT>(s, o, d, r, sk, cnt, overflow);
}

// Any hit in (1e-4, tmax) (occludedScene, bvh_traversal.cuh:98-133).  (An
// entry pushed here already starts before tmax, which does not change: with
// distance entries pop_live drops nothing.)
template <bool FAST, bool COUNT, class SK>
__device__ __forceinline__ bool trace_any_s_impl(const DevScene& s, F3 o, F3 d, const RayInv& r, float tmax,
                                                 const SK& sk, TraceCounters& cnt,
                                                 uint32_t& overflow) {
    using E = typename SK::Entry;
    constexpr int TOTAL = SK::kTotal;
    if (s.num_nodes == 0) return false;
    float4 rlo = gld(s.nodes), rhi = gld(s.nodes + 1);
    if (COUNT) cnt.nodes += 1;
    if (slab_v<FAST>(r, rlo, rhi, tmax) < 0.0f) return false;
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
            float4 alo = gld(c), ahi = gld(c + 1), blo = gld(c + 2), bhi = gld(c + 3);
            if (COUNT) cnt.nodes += 2;
            float ta = slab_v<FAST>(r, alo, ahi, tmax);
            float tb = slab_v<FAST>(r, blo, bhi, tmax);
            uint32_t da = __float_as_uint(alo.w), db = __float_as_uint(blo.w);
            if (ta >= 0.0f && tb >= 0.0f) {
                uint32_t near = ta <= tb ? da : db, far = ta <= tb ? db : da;
                if (top < TOTAL) sk.put(top++, make_entry<E>(far, fmaxf(ta, tb)));
                else overflow++;
                cur = near;
                continue;
            }
            if (ta >= 0.0f) { cur = da; continue; }
            if (tb >= 0.0f) { cur = db; continue; }
        }
        if (!pop_live(sk, top, tmax, cur)) break;
    }
    return false;
}

template <bool COUNT, class SK>
__device__ __forceinline__ bool trace_any_s(const DevScene& s, F3 o, F3 d, float tmax,
                                            const SK& sk, TraceCounters& cnt,
                                            uint32_t& overflow) {
    const RayInv r = make_ray_inv(o, d);
    if (finite_inv(r)) return trace_any_s_impl<true, COUNT>(s, o, d, r, tmax, sk, cnt, overflow);
    return trace_any_s_impl<false, COUNT>(s, o, d, r, tmax, sk, cnt, overflow);
}

// ---- 4-wide BVH with quantized child boxes --------------------------------
// Node layout (64 B): {P.xyz, S.x}, {S.y, S.z, qlo.x, qlo.y},
// {qlo.z, qhi.x, qhi.y, qhi.z}, {child words}; child i's plane on axis a is
// fma(byte i of q.a, S.a, P.a), a box that contains the child's exact box.
// Child words: wide-node index, kLeafFlag | first triangle, or kWideEmpty.
// Hit children are entered nearest first (entry distance, then slot), the
// others pushed farthest first.  The triangle tests and the slab arithmetic
// are the binary traversal's, on boxes that contain the binary ones, so the
// closest hit is the same triangle up to exact-distance ties.

constexpr uint32_t kWideEmpty = 0xFFFFFFFFu;

__device__ __forceinline__ float wide_plane(uint32_t bytes, int i, float s, float p) {
    return fmaf(static_cast<float>((bytes >> (8 * i)) & 0xFFu), s, p);
}

// Slab-tests the four children of wide node nd against (r, tmax); returns how
// many were hit and leaves their words in w[0..n) nearest first, with their
// entry distances' bits in k[0..n).
template <bool FAST>
__device__ __forceinline__ int wide_children(const float4* nd, const RayInv& r, float tmax, uint32_t (&w)[4],
                                             uint32_t (&k)[4]) {
    const float4 a = gld(nd), b = gld(nd + 1), c = gld(nd + 2), cw = gld(nd + 3);
    const uint32_t qlx = __float_as_uint(b.z), qly = __float_as_uint(b.w), qlz = __float_as_uint(c.x);
    const uint32_t qhx = __float_as_uint(c.y), qhy = __float_as_uint(c.z), qhz = __float_as_uint(c.w);
    w[0] = __float_as_uint(cw.x);
    w[1] = __float_as_uint(cw.y);
    w[2] = __float_as_uint(cw.z);
    w[3] = __float_as_uint(cw.w);
    // k: sort key, the entry distance's bits (t >= 0), or ~0 for a miss
    if constexpr (FAST) {
        // A plane at byte q lies at distance (q S + P - o) inv = q A + B with
        // A = S inv (exact: S is a power of two) and B = (P - o) inv, one fma
        // per plane.  The fma form and the plane-by-plane form
        // ((fma(q, S, P) - o) * inv, the decoded box under the reference's
        // slab arithmetic) round differently, by at most
        // u (1530 |A| + 5 |B| + 3 |P inv|) with u = 2^-24 (the decode's
        // rounding, the two roundings of B, the fma's, and those of the
        // plane-by-plane form).  With |P inv| <= |B| + |o inv| the slack e =
        // 2^-20 (255 |A| + 2 |B| + |o inv|) exceeds that: near planes move
        // closer by e and far planes farther, so every child box the
        // plane-by-plane test accepts is accepted here (a superset of leaves
        // is visited; the closest hit among them is the same triangle up to
        // exact-distance ties).  With inv >= 0 the low plane is the near one,
        // so each axis's near and far bytes are picked once per node.
        const bool px = r.inv.x >= 0.0f, py = r.inv.y >= 0.0f, pz = r.inv.z >= 0.0f;
        const uint32_t nx = px ? qlx : qhx, fx = px ? qhx : qlx;
        const uint32_t ny = py ? qly : qhy, fy = py ? qhy : qly;
        const uint32_t nz = pz ? qlz : qhz, fz = pz ? qhz : qlz;
        const float Ax = a.w * r.inv.x, Ay = b.x * r.inv.y, Az = b.y * r.inv.z;
        const float Bx = (a.x - r.o.x) * r.inv.x, By = (a.y - r.o.y) * r.inv.y, Bz = (a.z - r.o.z) * r.inv.z;
        const float ex = fmaf(255.0f * 0x1p-20f, fabsf(Ax), fmaf(0x1p-19f, fabsf(Bx), r.oi.x));
        const float ey = fmaf(255.0f * 0x1p-20f, fabsf(Ay), fmaf(0x1p-19f, fabsf(By), r.oi.y));
        const float ez = fmaf(255.0f * 0x1p-20f, fabsf(Az), fmaf(0x1p-19f, fabsf(Bz), r.oi.z));
        const float Bnx = Bx - ex, Bfx = Bx + ex, Bny = By - ey, Bfy = By + ey, Bnz = Bz - ez, Bfz = Bz + ez;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float tnx = fmaf(static_cast<float>((nx >> (8 * i)) & 0xFFu), Ax, Bnx);
            const float tfx = fmaf(static_cast<float>((fx >> (8 * i)) & 0xFFu), Ax, Bfx);
            const float tny = fmaf(static_cast<float>((ny >> (8 * i)) & 0xFFu), Ay, Bny);
            const float tfy = fmaf(static_cast<float>((fy >> (8 * i)) & 0xFFu), Ay, Bfy);
            const float tnz = fmaf(static_cast<float>((nz >> (8 * i)) & 0xFFu), Az, Bnz);
            const float tfz = fmaf(static_cast<float>((fz >> (8 * i)) & 0xFFu), Az, Bfz);
            const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, 0.0f));
            const float tf = fminf(fminf(tfx, tfy), fminf(tfz, tmax));
            k[i] = (w[i] != kWideEmpty && tn <= tf) ? (__float_as_uint(tn) & 0x7FFFFFFFu) : 0xFFFFFFFFu;
        }
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float4 lo = make_float4(wide_plane(qlx, i, a.w, a.x), wide_plane(qly, i, b.x, a.y),
                                          wide_plane(qlz, i, b.y, a.z), 0.0f);
            const float4 hi = make_float4(wide_plane(qhx, i, a.w, a.x), wide_plane(qhy, i, b.x, a.y),
                                          wide_plane(qhz, i, b.y, a.z), 0.0f);
            const float t = slab(r, lo, hi, tmax);
            k[i] = (w[i] != kWideEmpty && t >= 0.0f) ? (__float_as_uint(t) & 0x7FFFFFFFu) : 0xFFFFFFFFu;
        }
    }
    // sorting network on (key, word); equal keys keep the lower slot first
    auto cswap = [&](int i, int j) {
        const bool sw = k[j] < k[i];
        const uint32_t ki = k[i], wi = w[i];
        k[i] = sw ? k[j] : ki;
        w[i] = sw ? w[j] : wi;
        k[j] = sw ? ki : k[j];
        w[j] = sw ? wi : w[j];
    };
    cswap(0, 1);
    cswap(2, 3);
    cswap(0, 2);
    cswap(1, 3);
    cswap(1, 2);
    return (k[0] != 0xFFFFFFFFu) + (k[1] != 0xFFFFFFFFu) + (k[2] != 0xFFFFFFFFu) + (k[3] != 0xFFFFFFFFu);
}

// Pushes w[1..n) farthest first (with their entry distances k, for stacks
// that keep them); returns w[0].  The wide build keeps only trees whose
// stack need fits TOTAL entries, so no push can overflow.  While three
// entries fit below the LDS part's end, the three slots above top are
// written unconditionally (slots at and above top hold nothing live) and top
// advances by n - 1: one LDS address for every case.
template <class SK>
__device__ __forceinline__ uint32_t wide_descend(const uint32_t (&w)[4], const uint32_t (&k)[4], int n,
                                                 const SK& sk, int& top) {
    using E = typename SK::Entry;
    const int m = n - 1;
    const E e1 = make_entry<E>(w[1], __uint_as_float(k[1])), e2 = make_entry<E>(w[2], __uint_as_float(k[2])),
            e3 = make_entry<E>(w[3], __uint_as_float(k[3]));
    if (top + 3 <= SK::kLds) {
        // (values, not elements picked by index: an indexed pick would put w in scratch)
        const E v1 = m >= 3 ? e2 : e1;
        const E v0 = m >= 3 ? e3 : (m == 2 ? e2 : e1);
        sk.put3_lds(top, v0, v1, e1);
        top += m;
    } else {
        if (n > 3) sk.put(top++, e3);
        if (n > 2) sk.put(top++, e2);
        if (n > 1) sk.put(top++, e1);
    }
    return w[0];
}

// Closest hit over the wide BVH (node 0 is the root).
template <bool FAST, bool COUNT, class SK>
__device__ __forceinline__ Hit trace_closest_w_impl(const DevScene& s, F3 o, F3 d, const RayInv& r,
                                                    const SK& sk, TraceCounters& cnt,
                                                    uint32_t& overflow) {
    Hit h;
    h.t = kInf;
    h.u = h.v = 0.0f;
    h.tri = 0xFFFFFFFFu;
    if (s.num_wnodes == 0) return h;
    uint32_t cur = 0;
    int top = 0;
    while (true) {
        if (cur & kLeafFlag) {
            uint32_t i = cur & ~kLeafFlag;
            while (true) {
                float u, v;
                uint32_t word;
                float t = tri_test(o, d, s.tris + 3 * i, u, v, word);
                if (COUNT) cnt.tris += 1;
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
            uint32_t w[4], k[4];
            if (COUNT) cnt.nodes += 2;  // 64 B: two 32 B node records
            const int n = wide_children<FAST>(s.wnodes + 4 * cur, r, h.t, w, k);
            if (n > 0) {
                cur = wide_descend(w, k, n, sk, top);
                continue;
            }
        }
        if (!pop_live(sk, top, h.t, cur)) break;
    }
    return h;
}

template <bool COUNT, class SK>
__device__ __forceinline__ Hit trace_closest_w(const DevScene& s, F3 o, F3 d, const SK& sk,
                                               TraceCounters& cnt, uint32_t& overflow) {
    const RayInv r = make_ray_inv(o, d);
    if (finite_inv(r)) return trace_closest_w_impl<true, COUNT>(s, o, d, r, sk, cnt, overflow);
    return trace_closest_w_impl<false, COUNT>(s, o, d, r, sk, cnt, overflow);
}

// Any hit in (1e-4, tmax) over the wide BVH.
template <bool FAST, bool COUNT, class SK>
__device__ __forceinline__ bool trace_any_w_impl(const DevScene& s, F3 o, F3 d, const RayInv& r, float tmax,
                                                 const SK& sk, TraceCounters& cnt,
                                                 uint32_t& overflow) {
    if (s.num_wnodes == 0) return false;
    uint32_t cur = 0;
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
            uint32_t w[4], k[4];
            if (COUNT) cnt.nodes += 2;
            const int n = wide_children<FAST>(s.wnodes + 4 * cur, r, tmax, w, k);
            if (n > 0) {
                cur = wide_descend(w, k, n, sk, top);
                continue;
            }
        }
        if (!pop_live(sk, top, tmax, cur)) break;
    }
    return false;
}

template <bool COUNT, class SK>
__device__ __forceinline__ bool trace_any_w(const DevScene& s, F3 o, F3 d, float tmax,
                                            const SK& sk, TraceCounters& cnt,
                                            uint32_t& overflow) {
    const RayInv r = make_ray_inv(o, d);
    if (finite_inv(r)) return trace_any_w_impl<true, COUNT>(s, o, d, r, tmax, sk, cnt, overflow);
    return trace_any_w_impl<false, COUNT>(s, o, d, r, tmax, sk, cnt, overflow);
}

// Either BVH: the wide one when WIDE.
template <bool WIDE, bool COUNT, class SK>
__device__ __forceinline__ Hit trace_closest_v(const DevScene& s, F3 o, F3 d, const SK& sk,
                                               TraceCounters& cnt, uint32_t& overflow) {
    if constexpr (WIDE) return trace_closest_w<COUNT>(s, o, d, sk, cnt, overflow);
    else return trace_closest_s<COUNT>(s, o, d, sk, cnt, overflow);
}
template <bool WIDE, bool COUNT, class SK>
__device__ __forceinline__ bool trace_any_v(const DevScene& s, F3 o, F3 d, float tmax,
                                            const SK& sk, TraceCounters& cnt,
                                            uint32_t& overflow) {
    if constexpr (WIDE) return trace_any_w<COUNT>(s, o, d, tmax, sk, cnt, overflow);
    else return trace_any_s<COUNT>(s, o, d, tmax, sk, cnt, overflow);
}

}  // namespace ptd

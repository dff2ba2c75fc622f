// Small float3 algebra for gfx950 kernels.  Semantics follow the reference's
// vector library (Part_2/src/math/vecmath.h, math_utils.h): normalize divides
// by the length and maps a zero vector to zero; min/max are fminf/fmaxf
// (NaN-dropping).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ptd {

// No multiply-add contraction from here to the end of this file: every
// product and sum rounds as written, so each kernel that inlines these
// functions (the megakernel, the wavefront forward pass, the adjoint's
// replay) computes a path's radiance to the same bits, whatever code
// surrounds them.
#pragma clang fp contract(off)

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.31830988618379067154f;
constexpr float kInv2Pi = 0.15915494309189533577f;
constexpr float kTwoPi = 6.28318530717958647692f;
constexpr float kInf = 3.402823466e+38f;  // FLT_MAX, the reference's Infinity (pt.h:40)
constexpr float kEps = 1.192092896e-07f;  // FLT_EPSILON

struct F3 {
    float x, y, z;
};

__device__ __forceinline__ F3 f3(float x, float y, float z) { return F3{x, y, z}; }
__device__ __forceinline__ F3 f3(float s) { return F3{s, s, s}; }
__device__ __forceinline__ F3 f3(const float4& v) { return F3{v.x, v.y, v.z}; }
__device__ __forceinline__ F3 operator+(F3 a, F3 b) { return F3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ F3 operator-(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ F3 operator*(F3 a, F3 b) { return F3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ F3 operator*(F3 a, float s) { return F3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ F3 operator*(float s, F3 a) { return F3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ F3 operator/(F3 a, float s) { return F3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ F3 operator-(F3 a) { return F3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ F3& operator+=(F3& a, F3 b) { a = a + b; return a; }
__device__ __forceinline__ F3& operator*=(F3& a, F3 b) { a = a * b; return a; }
__device__ __forceinline__ F3& operator*=(F3& a, float s) { a = a * s; return a; }
__device__ __forceinline__ F3& operator/=(F3& a, float s) { a = a / s; return a; }
__device__ __forceinline__ float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ F3 cross(F3 a, F3 b) {
    return F3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float length(F3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ __forceinline__ F3 normalize(F3 a) {
    float l = length(a);
    return l > 0.0f ? a / l : F3{0.0f, 0.0f, 0.0f};
}
// Round-to-nearest forms without FMA contraction: every product and sum
// rounds separately, as the reference's host arithmetic does.  Used where
// results are compared bit for bit (triangle test, ray normalization).
__device__ __forceinline__ float dot_rn(F3 a, F3 b) {
#pragma clang fp contract(off)
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
__device__ __forceinline__ F3 cross_rn(F3 a, F3 b) {
#pragma clang fp contract(off)
    return F3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ F3 normalize_rn(F3 a) {
    float l = sqrtf(dot_rn(a, a));
    return l > 0.0f ? a / l : F3{0.0f, 0.0f, 0.0f};
}
__device__ __forceinline__ bool is_zero(F3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
__device__ __forceinline__ F3 reflect(F3 i, F3 n) { return i - 2.0f * dot(i, n) * n; }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__device__ __forceinline__ float max3(F3 a) { return fmaxf(fmaxf(a.x, a.y), a.z); }

// Orthonormal basis around n (math/frame.h:11-34).
struct Frame {
    F3 t, b, n;
};
__device__ __forceinline__ Frame make_frame(F3 n) {
    F3 helper = fabsf(n.x) > 0.999f ? f3(0.0f, 0.0f, 1.0f) : f3(1.0f, 0.0f, 0.0f);
    Frame f;
    f.n = n;
    f.t = normalize(cross(n, helper));
    f.b = normalize(cross(n, f.t));
    return f;
}
__device__ __forceinline__ F3 to_world(const Frame& f, F3 v) { return v.x * f.t + v.y * f.b + v.z * f.n; }
__device__ __forceinline__ F3 to_local(const Frame& f, F3 v) { return f3(dot(v, f.t), dot(v, f.b), dot(v, f.n)); }

// Wave-level sum for 64-wide wavefronts (used to aggregate counters).
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// (the translation unit's default again)
#pragma clang fp contract(fast)

}  // namespace ptd

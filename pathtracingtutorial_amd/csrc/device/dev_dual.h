// Forward-mode dual numbers for the BSDF evaluation template.
//
// Dual<N> carries a value and N partial derivatives; the helpers sval /
// smaxf / sminf / ssqrt / sabs of dev_bsdf.h get overloads, so
// eval_bsdf<Dual<N>> returns f together with its derivatives with respect to
// the seeded material parameters.  max/min pick a branch (derivative of the
// chosen operand), as automatic differentiation of a piecewise function does.
#pragma once

#include "dev_bsdf.h"

namespace ptd {

template <int N>
struct Dual {
    float v;
    float d[N];
    __device__ __forceinline__ Dual() {}
    __device__ __forceinline__ Dual(float x) : v(x) {
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] = 0.0f;
    }
    __device__ __forceinline__ static Dual seed(float x, int k) {
        Dual r(x);
        r.d[k] = 1.0f;
        return r;
    }
};

template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a.v + b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a.v - b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a) {
    Dual<N> r;
    r.v = -a.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a.v * b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    const float inv = 1.0f / b.v;
    r.v = a.v * inv;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * inv;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, float b) {
    Dual<N> r = a;
    r.v = a.v + b;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator+(float a, const Dual<N>& b) {
    return b + a;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, float b) {
    Dual<N> r = a;
    r.v = a.v - b;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(float a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a - b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = -b.d[i];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, float b) {
    Dual<N> r;
    r.v = a.v * b;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(float a, const Dual<N>& b) {
    return b * a;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, float b) {
    Dual<N> r;
    r.v = a.v / b;
    const float inv = 1.0f / b;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * inv;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(float a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a / b.v;
    const float k = -r.v / b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = k * b.d[i];
    return r;
}

template <int N>
__device__ __forceinline__ float sval(const Dual<N>& a) {
    return a.v;
}
template <int N>
__device__ __forceinline__ Dual<N> smaxf(const Dual<N>& a, const Dual<N>& b) {
    return a.v >= b.v ? a : b;
}
template <int N>
__device__ __forceinline__ Dual<N> smaxf(const Dual<N>& a, float b) {
    return a.v >= b ? a : Dual<N>(b);
}
template <int N>
__device__ __forceinline__ Dual<N> sminf(const Dual<N>& a, const Dual<N>& b) {
    return a.v <= b.v ? a : b;
}
template <int N>
__device__ __forceinline__ Dual<N> sminf(const Dual<N>& a, float b) {
    return a.v <= b ? a : Dual<N>(b);
}
template <int N>
__device__ __forceinline__ Dual<N> ssqrt(const Dual<N>& a) {
    Dual<N> r;
    r.v = sqrtf(a.v);
    const float k = r.v > 0.0f ? 0.5f / r.v : 0.0f;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * k;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> sabs(const Dual<N>& a) {
    return a.v < 0.0f ? -a : a;
}

// Partials carried through the adjoint's BSDF evaluations: the resolved
// base colour (3), roughness and metallic.
constexpr int kDP = 5;
constexpr int kDBase = 0, kDRough = 3, kDMetal = 4;
using D5 = Dual<kDP>;

// NP = kDP carries all five partials; NP = 3 only the base colour's (the
// first three), for batches that want no roughness or metallic gradient.
template <int NP = kDP>
__device__ __forceinline__ MatEval<Dual<NP>> to_dual(const MatEval<float>& m) {
    static_assert(NP == kDP || NP == 3, "partials: all five, or the base colour's");
    using D = Dual<NP>;
    MatEval<D> e;
    e.base = sv3(D::seed(m.base.x, kDBase + 0), D::seed(m.base.y, kDBase + 1), D::seed(m.base.z, kDBase + 2));
    if constexpr (NP == kDP) {
        e.metallic = D::seed(m.metallic, kDMetal);
        e.roughness = D::seed(m.roughness, kDRough);
    } else {
        e.metallic = D(m.metallic);
        e.roughness = D(m.roughness);
    }
    e.spec_trans = D(m.spec_trans);
    e.ior = D(m.ior);
    e.clearcoat = D(m.clearcoat);
    e.clearcoat_roughness = D(m.clearcoat_roughness);
    e.sheen = D(m.sheen);
    e.sheen_tint = D(m.sheen_tint);
    e.subsurface = D(m.subsurface);
    e.anisotropic = D(m.anisotropic);
    e.specular_tint = D(m.specular_tint);
    e.specular_factor = D(m.specular_factor);
    e.specular_color = sv3(D(m.specular_color.x), D(m.specular_color.y), D(m.specular_color.z));
    e.sheen_color = sv3(D(m.sheen_color.x), D(m.sheen_color.y), D(m.sheen_color.z));
    e.emissive = m.emissive;
    e.emissive_strength = m.emissive_strength;
    e.anisotropy_rotation = m.anisotropy_rotation;
    e.thickness = m.thickness;
    e.attenuation_distance = m.attenuation_distance;
    e.attenuation_color = m.attenuation_color;
    return e;
}

}  // namespace ptd

// Disney principled BSDF: material fetch, evaluation and sampling.
//
// Follows Part_2/src/materials/disney_bsdf.cuh and microfacet.cuh:
// Burley diffuse with the subsurface blend, sheen, anisotropic GGX with the
// height-correlated Smith G2 and a tinted Schlick F0, clearcoat (GGX, F0
// 0.04, weight 0.25), rough dielectric transmission, and lobe selection by
// the weights of computeLobeWeights.  Like the reference, evaluation uses the
// frame of N only (the tangent argument has no effect) and returns f*NoL for
// reflection and f for transmission, with the mixture pdf.
//
// Evaluation is a template over the scalar type S of the material parameters
// (float for rendering, a dual number for the adjoint); directions are float.
// The two uniform numbers of every 2-D warp are drawn in argument order
// (first draw -> first argument), written as separate statements.
#pragma once

#include "dev_math.h"
#include "dev_sampler.h"
#include "dev_scene.h"

namespace ptd {

// No multiply-add contraction from here to the end of this file: every
// product and sum rounds as written, so each kernel that inlines these
// functions (the megakernel, the wavefront forward pass, the adjoint's
// replay) computes a path's radiance to the same bits, whatever code
// surrounds them.
#pragma clang fp contract(off)

// ---- scalar helpers usable with float and with dual numbers --------------
__device__ __forceinline__ float sval(float x) { return x; }
__device__ __forceinline__ float smaxf(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ float sminf(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ float ssqrt(float a) { return sqrtf(a); }
__device__ __forceinline__ float sabs(float a) { return fabsf(a); }

template <class S>
struct SV3 {
    S x, y, z;
};
template <class S>
__device__ __forceinline__ SV3<S> sv3(S x, S y, S z) {
    return SV3<S>{x, y, z};
}

// Resolved material (MaterialEval, disney_bsdf.cuh:20-36).
template <class S>
struct MatEval {
    SV3<S> base;
    S metallic, roughness, spec_trans, ior;
    S clearcoat, clearcoat_roughness;
    S sheen, sheen_tint, subsurface;
    S anisotropic, specular_tint, specular_factor;
    SV3<S> specular_color, sheen_color;
    // not differentiated
    F3 emissive;
    float emissive_strength, anisotropy_rotation, thickness, attenuation_distance;
    F3 attenuation_color;
};

template <class S>
struct BsdfEvalT {
    SV3<S> f;
    S pdf;
};

struct BsdfSample {
    F3 wi;
    F3 weight;
    float pdf;
    bool delta, valid;
};

// ---- textures (bilinear, normalized coordinates, CUDA address modes) ------
// Filtering follows the documented texture unit behaviour: xB = u*W - 0.5,
// weights from frac(xB) held with 8 fractional bits, texel indices folded by
// the address mode (wrap / clamp / mirror).
__device__ __forceinline__ int fold_index(int i, int n, int mode) {
    if (mode == 1) return i < 0 ? 0 : (i >= n ? n - 1 : i);
    if (mode == 2) {
        int p = 2 * n;
        int m = i % p;
        if (m < 0) m += p;
        return m < n ? m : p - 1 - m;
    }
    int m = i % n;
    return m < 0 ? m + n : m;
}

__device__ __forceinline__ float filt_weight(float f) { return rintf(f * 256.0f) * (1.0f / 256.0f); }

// The four texels and weights of a bilinear lookup (order 00, 10, 01, 11).
struct TexTaps {
    uint32_t idx[4];
    float w[4];
};
// The texel coordinate u W - 0.5 rounds the product and the difference on
// their own (no multiply-add fusion): its fraction is quantized to 1/256, and
// a fused form can land one step away from the documented arithmetic.
__device__ __forceinline__ TexTaps tex_taps(int w, int h, int wrap_s, int wrap_t, float u, float v) {
#pragma clang fp contract(off)
    if (wrap_s == 0) u = u - floorf(u);
    if (wrap_t == 0) v = v - floorf(v);
    float x = u * static_cast<float>(w) - 0.5f, y = v * static_cast<float>(h) - 0.5f;
    float fx = floorf(x), fy = floorf(y);
    float a = filt_weight(x - fx), b = filt_weight(y - fy);
    int x0 = static_cast<int>(fx), y0 = static_cast<int>(fy);
    int xa = fold_index(x0, w, wrap_s), xb = fold_index(x0 + 1, w, wrap_s);
    int ya = fold_index(y0, h, wrap_t), yb = fold_index(y0 + 1, h, wrap_t);
    TexTaps t;
    t.idx[0] = static_cast<uint32_t>(ya * w + xa);
    t.idx[1] = static_cast<uint32_t>(ya * w + xb);
    t.idx[2] = static_cast<uint32_t>(yb * w + xa);
    t.idx[3] = static_cast<uint32_t>(yb * w + xb);
    t.w[0] = (1.0f - a) * (1.0f - b);
    t.w[1] = a * (1.0f - b);
    t.w[2] = (1.0f - a) * b;
    t.w[3] = a * b;
    return t;
}

// The weighted sum of a bilinear lookup, each product and sum rounded on its own.
__device__ __forceinline__ float4 texel_blend(float w00, float w10, float w01, float w11, float4 t00, float4 t10,
                                              float4 t01, float4 t11) {
#pragma clang fp contract(off)
    return make_float4(w00 * t00.x + w10 * t10.x + w01 * t01.x + w11 * t11.x,
                       w00 * t00.y + w10 * t10.y + w01 * t01.y + w11 * t11.y,
                       w00 * t00.z + w10 * t10.z + w01 * t01.z + w11 * t11.z,
                       w00 * t00.w + w10 * t10.w + w01 * t01.w + w11 * t11.w);
}

__device__ __forceinline__ float4 tex_bilinear(const float4* texels, int w, int h, int wrap_s, int wrap_t, float u,
                                               float v) {
    TexTaps tp = tex_taps(w, h, wrap_s, wrap_t, u, v);
    float4 t00 = texels[tp.idx[0]], t10 = texels[tp.idx[1]];
    float4 t01 = texels[tp.idx[2]], t11 = texels[tp.idx[3]];
    float w00 = tp.w[0], w10 = tp.w[1], w01 = tp.w[2], w11 = tp.w[3];
    return texel_blend(w00, w10, w01, w11, t00, t10, t01, t11);
}

__device__ __forceinline__ bool tex_valid(const DevScene& s, int id) {
    return id >= 0 && s.textures != nullptr && static_cast<uint32_t>(id) < s.num_textures;
}

__device__ __forceinline__ float4 tex_fetch(const DevScene& s, int id, float u, float v) {
    const TexDesc& t = s.textures[id];
    return tex_bilinear(s.texels + t.offset, t.w, t.h, t.wrap_s, t.wrap_t, u, v);
}

// pow(x, 2.2) colour decode of the reference's texture reads (disney_bsdf.cuh:81-83)
__device__ __forceinline__ F3 tex_color(const DevScene& s, int id, float u, float v) {
    float4 c = tex_fetch(s, id, u, v);
    if (s.textures[id].linear) return f3(c.x, c.y, c.z);
    return f3(powf(c.x, 2.2f), powf(c.y, 2.2f), powf(c.z, 2.2f));
}

// shading_utils.cuh:17-27
__device__ __forceinline__ void apply_uv_transform(const pt_material& m, float& u, float& v) {
    if (m.uv_rotation == 0.0f && m.uv_scale_x == 1.0f && m.uv_scale_y == 1.0f && m.uv_offset_x == 0.0f &&
        m.uv_offset_y == 0.0f)
        return;
    float cr = cosf(m.uv_rotation), sr = sinf(m.uv_rotation);
    float ru = cr * u + sr * v, rv = -sr * u + cr * v;
    u = ru * m.uv_scale_x + m.uv_offset_x;
    v = rv * m.uv_scale_y + m.uv_offset_y;
}

// fetchMaterial (disney_bsdf.cuh:53-122), float instantiation.
__device__ __forceinline__ MatEval<float> fetch_material(const DevScene& s, const pt_material& m, float u, float v) {
    MatEval<float> e;
    e.base = sv3(m.base_color.x, m.base_color.y, m.base_color.z);
    e.metallic = m.metallic;
    e.roughness = m.roughness;
    e.spec_trans = m.spec_trans;
    e.ior = m.ior;
    e.clearcoat = m.clearcoat;
    e.clearcoat_roughness = m.clearcoat_roughness;
    e.sheen = m.sheen;
    e.sheen_tint = m.sheen_tint;
    e.subsurface = m.subsurface;
    e.anisotropic = m.anisotropic;
    e.specular_tint = m.specular_tint;
    e.specular_factor = m.specular_factor;
    e.specular_color = sv3(m.specular_color_r, m.specular_color_g, m.specular_color_b);
    e.emissive = f3(m.emissive.x, m.emissive.y, m.emissive.z);
    e.emissive_strength = m.emissive_strength;
    e.sheen_color = sv3(m.sheen_color_r, m.sheen_color_g, m.sheen_color_b);
    e.anisotropy_rotation = m.anisotropy_rotation;
    e.thickness = m.thickness_factor;
    e.attenuation_distance = m.attenuation_distance;
    e.attenuation_color = f3(m.att_color_r, m.att_color_g, m.att_color_b);
    if (s.num_textures > 0 && s.textures != nullptr) {
        if (tex_valid(s, m.base_color_tex_id)) {
            F3 c = tex_color(s, m.base_color_tex_id, u, v);
            e.base = sv3(e.base.x * c.x, e.base.y * c.y, e.base.z * c.z);
        }
        if (tex_valid(s, m.metallic_roughness_tex_id)) {
            float4 mr = tex_fetch(s, m.metallic_roughness_tex_id, u, v);
            e.roughness *= mr.y;
            e.metallic *= mr.z;
        }
        if (tex_valid(s, m.emissive_tex_id)) e.emissive = e.emissive * tex_color(s, m.emissive_tex_id, u, v);
        if (tex_valid(s, m.clearcoat_tex_id)) e.clearcoat *= tex_fetch(s, m.clearcoat_tex_id, u, v).x;
        if (tex_valid(s, m.clearcoat_roughness_tex_id))
            e.clearcoat_roughness *= tex_fetch(s, m.clearcoat_roughness_tex_id, u, v).y;
        if (tex_valid(s, m.transmission_tex_id)) e.spec_trans *= tex_fetch(s, m.transmission_tex_id, u, v).x;
        if (tex_valid(s, m.sheen_color_tex_id)) {
            F3 c = tex_color(s, m.sheen_color_tex_id, u, v);
            e.sheen_color = sv3(e.sheen_color.x * c.x, e.sheen_color.y * c.y, e.sheen_color.z * c.z);
        }
        if (tex_valid(s, m.sheen_roughness_tex_id)) e.sheen *= tex_fetch(s, m.sheen_roughness_tex_id, u, v).w;
    }
    e.roughness = clampf(e.roughness, 0.001f, 1.0f);
    e.metallic = clampf(e.metallic, 0.0f, 1.0f);
    e.clearcoat_roughness = clampf(e.clearcoat_roughness, 0.001f, 1.0f);
    return e;
}

// ---- microfacet pieces (microfacet.cuh) -----------------------------------
template <class S>
__device__ __forceinline__ S d_ggx(float noh, S alpha) {  // :10-14
    S a2 = alpha * alpha;
    S d = noh * noh * (a2 - 1.0f) + 1.0f;
    return a2 / (kPi * d * d + 1e-7f);
}
template <class S>
__device__ __forceinline__ S d_ggx_aniso(float noh, float hx, float hy, S ax, S ay) {  // :17-22
    S x = hx / ax;
    S y = hy / ay;
    S d = x * x + y * y + noh * noh;
    return 1.0f / (kPi * ax * ay * d * d + 1e-7f);
}
template <class S>
__device__ __forceinline__ S g1_smith(float nov, S alpha) {  // :25-28
    S a2 = alpha * alpha;
    return 2.0f * nov / (nov + ssqrt(a2 + (1.0f - a2) * nov * nov) + 1e-7f);
}
template <class S>
__device__ __forceinline__ S g2_smith(float nol, float nov, S alpha) {  // :31-36
    S a2 = alpha * alpha;
    S gv = nol * ssqrt(a2 + (1.0f - a2) * nov * nov);
    S gl = nov * ssqrt(a2 + (1.0f - a2) * nol * nol);
    return 2.0f * nol * nov / (gv + gl + 1e-7f);
}
// x^5 by multiplication (the reference writes powf(x, 5.0f); the two agree
// to a few ulp and the product needs no log/exp)
__device__ __forceinline__ float pow5(float x) {
    float x2 = x * x;
    return x2 * x2 * x;
}
__device__ __forceinline__ float schlick_t5(float cos_theta) {  // :39-51
    float t = clampf(1.0f - cos_theta, 0.0f, 1.0f);
    float t2 = t * t;
    return t2 * t2 * t;
}
// scalar Schlick keeps its own association: F0 + (((1-F0)*t2)*t2)*t
__device__ __forceinline__ float schlick_scalar(float cos_theta, float f0) {
    float t = clampf(1.0f - cos_theta, 0.0f, 1.0f);
    float t2 = t * t;
    return f0 + (1.0f - f0) * t2 * t2 * t;
}
template <class S>
__device__ __forceinline__ S f_dielectric(float cos_i, S eta) {  // :54-62
    S sin2t = eta * eta * (1.0f - cos_i * cos_i);
    if (sval(sin2t) >= 1.0f) return S(1.0f);
    S cos_t = ssqrt(smaxf(S(0.0f), 1.0f - sin2t));
    float ci = fabsf(cos_i);
    S rs = (ci - eta * cos_t) / (ci + eta * cos_t + 1e-7f);
    S rp = (eta * ci - cos_t) / (eta * ci + cos_t + 1e-7f);
    return 0.5f * (rs * rs + rp * rp);
}

// VNDF sampling (microfacet.cuh:67-92), float only.
__device__ __forceinline__ F3 sample_vndf(F3 wo, float ax, float ay, float u1, float u2) {
    F3 wh = normalize(f3(ax * wo.x, ay * wo.y, wo.z));
    if (wh.z < 0.0f) wh = -wh;
    F3 t1v = (wh.z < 0.9999f) ? normalize(cross(f3(0.0f, 0.0f, 1.0f), wh)) : f3(1.0f, 0.0f, 0.0f);
    F3 t2v = cross(wh, t1v);
    float r = sqrtf(u1);
    float phi = kTwoPi * u2;
    float t1 = r * cosf(phi);
    float t2 = r * sinf(phi);
    float sv = 0.5f * (1.0f + wh.z);
    t2 = (1.0f - sv) * sqrtf(fmaxf(0.0f, 1.0f - t1 * t1)) + sv * t2;
    F3 nh = t1 * t1v + t2 * t2v + sqrtf(fmaxf(0.0f, 1.0f - t1 * t1 - t2 * t2)) * wh;
    return normalize(f3(ax * nh.x, ay * nh.y, fmaxf(0.0f, nh.z)));
}

template <class S>
__device__ __forceinline__ S pdf_vndf(F3 wo, F3 h, S ax, S ay) {  // :95-102
    float nov = fmaxf(wo.z, 1e-5f);
    float noh = fmaxf(h.z, 0.0f);
    float voh = fmaxf(dot(wo, h), 0.0f);
    S d = d_ggx_aniso(noh, h.x, h.y, ax, ay);
    S g1 = g1_smith(nov, ssqrt(ax * ay));
    return d * g1 * voh / (nov + 1e-7f);
}

__device__ __forceinline__ bool refract_dir(F3 wi, F3 n, float eta, F3& wt) {  // :105-113
    float ci = dot(wi, n);
    float s2i = fmaxf(0.0f, 1.0f - ci * ci);
    float s2t = eta * eta * s2i;
    if (s2t >= 1.0f) return false;
    float ct = sqrtf(1.0f - s2t);
    wt = eta * (-wi) + (eta * ci - ct) * n;
    return true;
}

// cosine hemisphere (math/sampling.h:18-25)
__device__ __forceinline__ F3 sample_cosine(float u1, float u2) {
    float r = sqrtf(u1);
    float phi = kTwoPi * u2;
    return f3(r * cosf(phi), r * sinf(phi), sqrtf(fmaxf(0.0f, 1.0f - u1)));
}

__device__ __forceinline__ float power_heuristic(float a, float b) {  // sampling.h:34-39
    float a2 = a * a, b2 = b * b, sum = a2 + b2;
    return sum > 0.0f ? a2 / sum : 0.0f;
}

// ---- Disney lobes -----------------------------------------------------------
template <class S>
__device__ __forceinline__ void alpha_xy(S rough, S aniso, S& ax, S& ay) {  // disney_bsdf.cuh:126-131
    S aspect = ssqrt(1.0f - 0.9f * aniso);
    S a = rough * rough;
    ax = smaxf(a / aspect, 0.001f);
    ay = smaxf(a * aspect, 0.001f);
}

template <class S>
__device__ __forceinline__ S lum3(const SV3<S>& c) {
    return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z;
}

template <class S>
__device__ __forceinline__ SV3<S> tint_of(const SV3<S>& base) {
    S l = lum3(base);
    if (sval(l) > 0.0f) return sv3(base.x / l, base.y / l, base.z / l);
    return sv3(S(1.0f), S(1.0f), S(1.0f));
}

template <class S>
__device__ __forceinline__ SV3<S> compute_f0(const MatEval<S>& m) {  // :175-183
    SV3<S> tint = tint_of(m.base);
    S eta = m.ior;
    S r0 = ((eta - 1.0f) * (eta - 1.0f)) / ((eta + 1.0f) * (eta + 1.0f));
    S one_m = 1.0f - m.specular_tint;
    SV3<S> spec = sv3(r0 * (one_m * 1.0f + m.specular_tint * tint.x), r0 * (one_m * 1.0f + m.specular_tint * tint.y),
                      r0 * (one_m * 1.0f + m.specular_tint * tint.z));
    spec = sv3(spec.x * m.specular_factor * m.specular_color.x, spec.y * m.specular_factor * m.specular_color.y,
               spec.z * m.specular_factor * m.specular_color.z);
    S om = 1.0f - m.metallic;
    return sv3(om * spec.x + m.metallic * m.base.x, om * spec.y + m.metallic * m.base.y,
               om * spec.z + m.metallic * m.base.z);
}

template <class S>
struct LobeW {
    S diff, spec, cc, trans, sheen;
};

template <class S>
__device__ __forceinline__ LobeW<S> lobe_weights(const MatEval<S>& m) {  // :187-205
    LobeW<S> w;
    S dw = (1.0f - m.metallic) * (1.0f - m.spec_trans);
    S sw = S(1.0f);
    S tw = (1.0f - m.metallic) * m.spec_trans;
    S cw = 0.25f * m.clearcoat;
    S shw = (1.0f - m.metallic) * m.sheen * 0.1f;
    S total = dw + sw + cw + tw + shw;
    if (sval(total) <= 0.0f) {
        w.diff = w.spec = w.cc = w.trans = w.sheen = S(0.0f);
        return w;
    }
    S inv = 1.0f / total;
    w.diff = dw * inv;
    w.spec = sw * inv;
    w.cc = cw * inv;
    w.trans = tw * inv;
    w.sheen = shw * inv;
    return w;
}

// evalDisneyBsdf (disney_bsdf.cuh:209-321).
template <class S>
__device__ __forceinline__ BsdfEvalT<S> eval_bsdf(const MatEval<S>& m, F3 wo, F3 wi, F3 n, bool entering) {
    BsdfEvalT<S> r;
    r.f = sv3(S(0.0f), S(0.0f), S(0.0f));
    r.pdf = S(0.0f);
    float nol = dot(n, wi);
    float nov = dot(n, wo);
    bool refl = nol > 0.0f;
    if (!refl && sval(m.spec_trans) <= 0.0f) return r;
    if (!refl) {
        float anl = fabsf(nol), anv = fabsf(nov);
        if (anl < 1e-5f || anv < 1e-5f) return r;
        S eta = entering ? (1.0f / m.ior) : m.ior;
        float etaf = sval(eta);
        // the half vector is a direction: computed at the primal eta
        F3 ht = -normalize(wo * etaf + wi);
        float hov = fabsf(dot(ht, wo));
        S ax, ay;
        alpha_xy(m.roughness, m.anisotropic, ax, ay);
        Frame fr = make_frame(n);
        F3 hl = to_local(fr, ht);
        S D = d_ggx_aniso(fabsf(hl.z), hl.x, hl.y, ax, ay);
        S G = g2_smith(anl, anv, ssqrt(ax * ay));
        S F = f_dielectric(hov, eta);
        S sd = dot(wo, ht) * eta + dot(wi, ht);
        S factor = fabsf(dot(wo, ht)) * fabsf(dot(wi, ht)) / (anv * anl * sd * sd + 1e-7f);
        S tw = (1.0f - m.metallic) * m.spec_trans;
        S omf = 1.0f - F;
        r.f = sv3(m.base.x * tw * D * G * omf * factor, m.base.y * tw * D * G * omf * factor,
                  m.base.z * tw * D * G * omf * factor);
        LobeW<S> w = lobe_weights(m);
        S jac = eta * eta * fabsf(dot(wi, ht)) / (sd * sd + 1e-7f);
        S tpdf = D * g1_smith(anv, ssqrt(ax * ay)) * hov / (anv + 1e-7f) * jac;
        r.pdf = w.trans * tpdf;
        return r;
    }
    if (nol < 1e-5f || nov < 1e-5f) return r;
    F3 H = normalize(wo + wi);
    float noh = fmaxf(dot(n, H), 0.0f);
    float hol = fmaxf(dot(H, wi), 0.0f);
    Frame fr = make_frame(n);
    F3 hl = to_local(fr, H);
    S ax, ay;
    alpha_xy(m.roughness, m.anisotropic, ax, ay);
    LobeW<S> w = lobe_weights(m);
    SV3<S> ft = sv3(S(0.0f), S(0.0f), S(0.0f));
    S pdf = S(0.0f);
    S diff_w = (1.0f - m.metallic) * (1.0f - m.spec_trans);
    if (sval(diff_w) > 0.0f) {  // evalDisneyDiffuse, :135-159
        S fd90 = 0.5f + 2.0f * hol * hol * m.roughness;
        float pl = pow5(1.0f - nol), pv = pow5(1.0f - nov);
        S fl = 1.0f + (fd90 - 1.0f) * pl;
        S fv = 1.0f + (fd90 - 1.0f) * pv;
        S dk = kInvPi * fl * fv;
        SV3<S> d = sv3(m.base.x * dk, m.base.y * dk, m.base.z * dk);
        if (sval(m.subsurface) > 0.0f) {
            S fss90 = hol * hol * m.roughness;
            S fsl = 1.0f + (fss90 - 1.0f) * pl;
            S fsv = 1.0f + (fss90 - 1.0f) * pv;
            S ss = 1.25f * (fsl * fsv * (1.0f / (nol + nov + 1e-5f) - 0.5f) + 0.5f);
            S sk = kInvPi * ss;
            S om = 1.0f - m.subsurface;
            d = sv3(om * d.x + m.subsurface * (m.base.x * sk), om * d.y + m.subsurface * (m.base.y * sk),
                    om * d.z + m.subsurface * (m.base.z * sk));
        }
        ft = sv3(ft.x + diff_w * d.x, ft.y + diff_w * d.y, ft.z + diff_w * d.z);
        pdf = pdf + w.diff * (nol * kInvPi);
    }
    if (sval(m.sheen) > 0.0f && sval(m.metallic) < 1.0f) {  // evalSheen, :163-171
        SV3<S> tint = tint_of(m.base);
        S ot = 1.0f - m.sheen_tint;
        SV3<S> bs = sv3(ot * 1.0f + m.sheen_tint * tint.x, ot * 1.0f + m.sheen_tint * tint.y,
                        ot * 1.0f + m.sheen_tint * tint.z);
        float fh = pow5(1.0f - hol);
        S om = 1.0f - m.metallic;
        ft = sv3(ft.x + om * (m.sheen * (bs.x * m.sheen_color.x) * fh),
                 ft.y + om * (m.sheen * (bs.y * m.sheen_color.y) * fh),
                 ft.z + om * (m.sheen * (bs.z * m.sheen_color.z) * fh));
        pdf = pdf + w.sheen * (nol * kInvPi);
    }
    {   // GGX specular
        S D = d_ggx_aniso(noh, hl.x, hl.y, ax, ay);
        S G = g2_smith(nol, nov, ssqrt(ax * ay));
        SV3<S> f0 = compute_f0(m);
        float t5 = schlick_t5(hol);
        SV3<S> F = sv3(f0.x + (1.0f - f0.x) * t5, f0.y + (1.0f - f0.y) * t5, f0.z + (1.0f - f0.z) * t5);
        S dg = D * G;
        float den = 4.0f * nol * nov + 1e-7f;
        ft = sv3(ft.x + dg * F.x / den, ft.y + dg * F.y / den, ft.z + dg * F.z / den);
        F3 wol = to_local(fr, wo);
        S sp = pdf_vndf(wol, hl, ax, ay) / (4.0f * hol + 1e-7f);
        pdf = pdf + w.spec * sp;
    }
    if (sval(m.clearcoat) > 0.0f) {
        S ca = smaxf(m.clearcoat_roughness * m.clearcoat_roughness, 0.001f);
        S Dc = d_ggx(noh, ca);
        S Gc = g2_smith(nol, nov, ca);
        float Fc = schlick_scalar(hol, 0.04f);
        S cs = 0.25f * m.clearcoat * Dc * Gc * Fc / (4.0f * nol * nov + 1e-7f);
        ft = sv3(ft.x + cs, ft.y + cs, ft.z + cs);
        S cp = Dc * noh / (4.0f * hol + 1e-7f);
        pdf = pdf + w.cc * cp;
    }
    r.f = sv3(ft.x * nol, ft.y * nol, ft.z * nol);
    r.pdf = pdf;
    return r;
}

// sampleDisneyBsdf (disney_bsdf.cuh:325-410), float.
template <class Rng>
__device__ __forceinline__ BsdfSample sample_bsdf(const MatEval<float>& m, F3 wo, F3 n, Rng& rng, bool entering) {
    BsdfSample res;
    res.wi = f3(0.0f);
    res.weight = f3(0.0f);
    res.pdf = 0.0f;
    res.delta = false;
    res.valid = false;
    float nov = dot(n, wo);
    if (nov < 1e-5f) return res;
    LobeW<float> w = lobe_weights(m);
    float choice = rng.next();
    Frame fr = make_frame(n);
    F3 wol = to_local(fr, wo);
    float ax, ay;
    alpha_xy(m.roughness, m.anisotropic, ax, ay);
    F3 wi;
    bool is_delta = false;
    if (choice < w.diff + w.sheen) {
        float u1 = rng.next();
        float u2 = rng.next();
        wi = to_world(fr, sample_cosine(u1, u2));
    } else if (choice < w.diff + w.sheen + w.spec) {
        float u1 = rng.next();
        float u2 = rng.next();
        F3 H = to_world(fr, sample_vndf(wol, ax, ay, u1, u2));
        wi = reflect(-wo, H);
        if (dot(wi, n) <= 0.0f) return res;
        is_delta = m.roughness < 0.01f;
    } else if (choice < w.diff + w.sheen + w.spec + w.cc) {
        float ca = fmaxf(m.clearcoat_roughness * m.clearcoat_roughness, 0.001f);
        float u1 = rng.next();
        float u2 = rng.next();
        F3 H = to_world(fr, sample_vndf(wol, ca, ca, u1, u2));
        wi = reflect(-wo, H);
        if (dot(wi, n) <= 0.0f) return res;
    } else {
        float u1 = rng.next();
        float u2 = rng.next();
        F3 H = to_world(fr, sample_vndf(wol, ax, ay, u1, u2));
        float eta = entering ? (1.0f / m.ior) : m.ior;
        F3 wt;
        if (!refract_dir(wo, H, eta, wt)) {
            wi = reflect(-wo, H);
            if (dot(wi, n) <= 0.0f) return res;
        } else {
            wi = normalize(wt);
            is_delta = m.roughness < 0.01f;
        }
    }
    BsdfEvalT<float> ev = eval_bsdf(m, wo, wi, n, entering);
    if (ev.pdf <= 0.0f) return res;
    res.wi = wi;
    res.pdf = ev.pdf;
    res.delta = is_delta;
    res.valid = true;
    if (is_delta) {
        if (dot(n, wi) < 0.0f) {
            res.weight = f3(m.base.x, m.base.y, m.base.z);
        } else {
            SV3<float> f0 = compute_f0(m);
            F3 H = normalize(wo + wi);
            float t5 = schlick_t5(fmaxf(dot(H, wi), 0.0f));
            res.weight = f3(f0.x + (1.0f - f0.x) * t5, f0.y + (1.0f - f0.y) * t5, f0.z + (1.0f - f0.z) * t5);
        }
    } else {
        res.weight = f3(ev.f.x, ev.f.y, ev.f.z) / (ev.pdf + 1e-7f);
    }
    return res;
}

// (the translation unit's default again)
#pragma clang fp contract(fast)

}  // namespace ptd

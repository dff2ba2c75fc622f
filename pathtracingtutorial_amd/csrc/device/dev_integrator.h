// Unidirectional path tracer with next-event estimation and MIS.
//
// Follows Part_2/src/integrators/path.cu:31-206 and direct_lighting.cuh:
// camera ray with jitter (and thin lens), closest hit, environment on a miss
// (radiance = bilinear texel * 2 pi, MIS against the env pdf after a
// non-delta bounce), emitter hit ends the path (MIS against the triangle
// pdf), alpha mask / stochastic blend, NEE from an alias-sampled emissive
// triangle and an alias-sampled env texel (both when both exist), punctual
// lights without MIS, Disney BSDF sampling, offset 1e-4 along the geometric
// normal, Beer-Lambert, Russian roulette after depth 2 with p = min(max T,
// 0.95), and a per-channel clamp of the path radiance at 100.
#pragma once

#include "dev_bsdf.h"
#include "dev_traverse.h"

namespace ptd {

// No multiply-add contraction from here to the end of this file: every
// product and sum rounds as written, so each kernel that inlines these
// functions (the megakernel, the wavefront forward pass, the adjoint's
// replay) computes a path's radiance to the same bits, whatever code
// surrounds them.
#pragma clang fp contract(off)

// A per-path integrator's traversal context: the thread's stack, the tree it
// walks (the 4-wide BVH when WIDE) and its ray counters.
template <bool WIDE, class SK>
struct TraceCtxT {
    static constexpr bool kWide = WIDE;
    SK sk;
    TraceCounters cnt;   // closest-hit queries
    TraceCounters scnt;  // shadow (any-hit) queries
    uint32_t overflow;
    uint32_t closest, shadow;
};
// The binary tree with a 32-entry stack all in LDS.
using TraceCtx = TraceCtxT<false, TraceStack<kStackSize, kStackSize, uint32_t>>;

// ---- environment --------------------------------------------------------
// envmap.h:15-24 (each step rounded on its own: the texel fraction of u and v
// is quantized to 1/256 by the lookup)
__device__ __forceinline__ void dir_to_uv(F3 d, float& u, float& v) {
#pragma clang fp contract(off)
    u = atan2f(d.z, d.x);
    v = asinf(d.y);
    u *= kInv2Pi;
    v *= kInvPi;
    u += 0.5f;
    v += 0.5f;
    v = 1.0f - v;
}
// envmap.h:27-32
__device__ __forceinline__ F3 uv_to_dir(float u, float v) {
    float phi = (u - 0.5f) * kTwoPi;
    float theta = v * kPi;
    float st = sinf(theta);
    return normalize(f3(cosf(phi) * st, cosf(theta), sinf(phi) * st));
}
__device__ __forceinline__ bool has_env_texture(const DevScene& s) { return s.env != nullptr; }
// sampleEnvironmentRadiance (envmap.h:34-45): bilinear lookup times 2 pi, at
// the direction's equirect coordinates (u, v) = dir_to_uv(d).
__device__ __forceinline__ F3 env_radiance_uv(const DevScene& s, float u, float v) {
    if (s.env == nullptr) return f3(0.0f);
    float4 c = tex_bilinear(s.env, s.env_w, s.env_h, s.env_wrap_s, s.env_wrap_t, u, v);
    return f3(c.x, c.y, c.z) * kTwoPi;
}
__device__ __forceinline__ F3 env_radiance(const DevScene& s, F3 d) {
    if (s.env == nullptr) return f3(0.0f);
    float u, v;
    dir_to_uv(d, u, v);
    return env_radiance_uv(s, u, v);
}
// environmentPdf (light_pdf.cuh:12-27) of direction d with (u, v) = dir_to_uv(d)
__device__ __forceinline__ float env_pdf_uv(const DevScene& s, F3 d, float u, float v) {
    if (s.env == nullptr || s.env_pmf == nullptr || s.num_env_texels == 0 || s.env_w <= 0 || s.env_h <= 0) return 0.0f;
    int x = static_cast<int>(clampf(u * s.env_w, 0.0f, static_cast<float>(s.env_w - 1)));
    int y = static_cast<int>(clampf(v * s.env_h, 0.0f, static_cast<float>(s.env_h - 1)));
    float st = sqrtf(fmaxf(0.0f, 1.0f - d.y * d.y));
    if (st <= 1e-5f) return 0.0f;
    int idx = y * s.env_w + x;
    return s.env_pmf[idx] * static_cast<float>(s.env_w * s.env_h) / (kTwoPi * kPi * st);
}
__device__ __forceinline__ float env_pdf(const DevScene& s, F3 d) {
    float u, v;
    dir_to_uv(d, u, v);
    return env_pdf_uv(s, d, u, v);
}

// ---- lights ---------------------------------------------------------------
struct LightSample {
    F3 wi, radiance;
    float pdf;
    bool valid;
};

// direct_lighting.cuh:55-59
__device__ __forceinline__ uint32_t sample_alias(const float2* table, uint32_t n, float u1, float u2) {
    uint32_t idx = static_cast<uint32_t>(u1 * n);
    if (idx >= n) idx = n - 1;
    float2 e = table[idx];
    return u2 < e.x ? idx : __float_as_uint(e.y);
}

// Shadow ray of isOccluded (direct_lighting.cuh:36-51): origin pushed 1e-4
// along the normal towards dir, tmax = distance - 1e-3 (or unbounded).
struct ShadowRay {
    F3 o, d;
    float tmax;
};
__device__ __forceinline__ ShadowRay make_shadow(F3 p, F3 n, F3 dir, float max_dist) {
    ShadowRay r;
    float sign = dot(dir, n) >= 0.0f ? 1.0f : -1.0f;
    r.o = p + n * (1e-4f * sign);
    r.d = normalize(dir);
    r.tmax = (max_dist < kInf) ? (max_dist - 1e-3f) : kInf;
    return r;
}

template <bool COUNT, class Ctx>
__device__ __forceinline__ bool occluded(const DevScene& s, const ShadowRay& r, Ctx& c) {
    c.shadow++;
    return trace_any_v<Ctx::kWide, COUNT>(s, r.o, r.d, r.tmax, c.sk, c.scnt, c.overflow);
}

// sampleAreaLight (direct_lighting.cuh:75-121) up to, not including, the
// visibility test, which is its last rejection; sh gets the shadow ray.
// The half of sampleAreaLight (direct_lighting.cuh:75-121) that does not
// depend on the shading point: the emitter picked by the alias table, the
// point and normal on it, its area, pmf and radiance.  It consumes the four
// dimensions the reference draws (a1, a2, t1, t2) before any early exit.
struct AreaGeom {
    F3 lp, ln, radiance;
    float area, pmf;
    uint32_t key;  // sampler position the sample started at
};
__device__ __forceinline__ void area_geom(const DevScene& s, Sampler& rng, AreaGeom& g) {
    g.key = rng.key();
    float a1 = rng.next();
    float a2 = rng.next();
    uint4 ref = s.emissive[sample_alias(s.tri_alias, s.num_emissive, a1, a2)];
    uint4 face = s.faces[ref.x];
    float4 q0 = s.vpos[face.x], q1 = s.vpos[face.y], q2 = s.vpos[face.z];
    F3 p0 = f3(q0), p1 = f3(q1), p2 = f3(q2);
    float t1 = rng.next();
    float t2 = rng.next();
    // light_sampler.h:58-71
    float su0 = sqrtf(t1);
    float b0 = 1.0f - su0, b1 = t2 * su0, b2 = 1.0f - b0 - b1;
    g.lp = p0 * b0 + p1 * b1 + p2 * b2;
    F3 ln = normalize(f3(s.vnrm[face.x]) * b0 + f3(s.vnrm[face.y]) * b1 + f3(s.vnrm[face.z]) * b2);
    F3 e1 = p1 - p0, e2 = p2 - p0;
    if (length(ln) == 0.0f) ln = normalize(cross(e1, e2));
    g.ln = ln;
    g.area = 0.5f * length(cross(e1, e2));
    g.pmf = __uint_as_float(ref.y);
    const pt_material& lm = s.materials[face.w];
    g.radiance = f3(lm.emissive.x, lm.emissive.y, lm.emissive.z) * lm.emissive_strength;
}
// The rest of sampleAreaLight, at shading point p with normal n.
__device__ __forceinline__ LightSample area_finish(F3 p, F3 n, const AreaGeom& g, ShadowRay& sh) {
    LightSample ls;
    ls.valid = false;
    F3 tl = g.lp - p;
    float d2 = dot(tl, tl);
    if (d2 <= 1e-8f) return ls;
    float dist = sqrtf(d2);
    F3 wi = tl / dist;
    float cs = fmaxf(0.0f, dot(n, wi));
    if (cs <= 0.0f) return ls;
    if (g.area <= 0.0f) return ls;
    float cl = fabsf(dot(g.ln, -wi));
    if (cl <= 1e-5f) return ls;
    float pdf = (g.pmf / g.area) * d2 / cl;
    if (pdf <= 0.0f) return ls;
    sh = make_shadow(p, n, wi, dist);
    ls.wi = wi;
    ls.radiance = g.radiance;
    ls.pdf = pdf;
    ls.valid = true;
    return ls;
}
__device__ __forceinline__ LightSample sample_area_nv(const DevScene& s, F3 p, F3 n, Sampler& rng, ShadowRay& sh) {
    if (s.num_emissive == 0 || s.emissive == nullptr) {
        LightSample ls;
        ls.valid = false;
        return ls;
    }
    AreaGeom g;
    area_geom(s, rng, g);
    return area_finish(p, n, g, sh);
}

template <bool COUNT, class Ctx>
__device__ __forceinline__ LightSample sample_area(const DevScene& s, F3 p, F3 n, Sampler& rng, Ctx& c) {
    ShadowRay sh;
    LightSample ls = sample_area_nv(s, p, n, rng, sh);
    if (ls.valid && occluded<COUNT>(s, sh, c)) ls.valid = false;
    return ls;
}

// sampleEnvironmentLight (direct_lighting.cuh:123-154) without the final
// visibility test; sh gets the shadow ray.
// The half of sampleEnvLight (direct_lighting.cuh:123-154) that does not
// depend on the shading point: the texel picked by the alias table, the
// jittered direction in it, its pdf and radiance (four dimensions).
struct EnvGeom {
    F3 wi, radiance;
    float pdf;
    uint32_t key;  // sampler position the sample started at
};
__device__ __forceinline__ void env_geom(const DevScene& s, Sampler& rng, EnvGeom& g) {
    g.key = rng.key();
    float a1 = rng.next();
    float a2 = rng.next();
    uint32_t texel = sample_alias(s.env_alias, s.num_env_texels, a1, a2);
    int x = static_cast<int>(texel % static_cast<uint32_t>(s.env_w));
    int y = static_cast<int>(texel / static_cast<uint32_t>(s.env_w));
    float j1 = rng.next();
    float j2 = rng.next();
    float u = (static_cast<float>(x) + j1) / static_cast<float>(s.env_w);
    float v = (static_cast<float>(y) + j2) / static_cast<float>(s.env_h);
    g.wi = uv_to_dir(u, v);
    float wu, wv;  // the equirect coordinates of wi, shared by the pdf and the radiance lookup
    dir_to_uv(g.wi, wu, wv);
    g.pdf = env_pdf_uv(s, g.wi, wu, wv);
    g.radiance = env_radiance_uv(s, wu, wv);
}
// The rest of sampleEnvLight, at shading point p with normal n.
__device__ __forceinline__ LightSample env_finish(F3 p, F3 n, const EnvGeom& g, ShadowRay& sh) {
    LightSample ls;
    ls.valid = false;
    float cs = fmaxf(0.0f, dot(n, g.wi));
    if (cs <= 0.0f) return ls;
    if (g.pdf <= 0.0f) return ls;
    sh = make_shadow(p, n, g.wi, kInf);
    ls.wi = g.wi;
    ls.radiance = g.radiance;
    ls.pdf = g.pdf;
    ls.valid = true;
    return ls;
}
__device__ __forceinline__ LightSample sample_env_nv(const DevScene& s, F3 p, F3 n, Sampler& rng, ShadowRay& sh) {
    EnvGeom g;
    env_geom(s, rng, g);
    return env_finish(p, n, g, sh);
}

template <bool COUNT, class Ctx>
__device__ __forceinline__ LightSample sample_env(const DevScene& s, F3 p, F3 n, Sampler& rng, Ctx& c) {
    ShadowRay sh;
    LightSample ls = sample_env_nv(s, p, n, rng, sh);
    if (ls.valid && occluded<COUNT>(s, sh, c)) ls.valid = false;
    return ls;
}

// evaluateDirectLight (direct_lighting.cuh:204-218)
__device__ __forceinline__ F3 eval_direct(const LightSample& ls, const MatEval<float>& m, F3 wo, F3 n) {
    if (!ls.valid) return f3(0.0f);
    BsdfEvalT<float> b = eval_bsdf(m, wo, ls.wi, n, true);
    if (b.pdf <= 0.0f) return f3(0.0f);
    float mis = power_heuristic(ls.pdf, b.pdf);
    return ls.radiance * f3(b.f.x, b.f.y, b.f.z) * (mis / (ls.pdf + 1e-7f));
}

// One punctual light of evaluatePunctualLights (direct_lighting.cuh:158-200)
// up to its visibility test: direction, unoccluded intensity and distance;
// false when the light cannot contribute (too close, behind the surface).
__device__ __forceinline__ bool punctual_light_sample(const DevScene& s, uint32_t i, F3 p, F3 n, F3& wi, F3& inten,
                                                      float& dist) {
    const pt_punctual_light& l = s.lights[i];
    dist = kInf;
    if (l.type == 0 || l.type == 2) {
        F3 tl = f3(l.position.x, l.position.y, l.position.z) - p;
        float d2 = dot(tl, tl);
        if (d2 < 1e-8f) return false;
        dist = sqrtf(d2);
        wi = tl / dist;
        float att = dist >= l.range ? 0.0f : 1.0f / (dist * dist + 1e-7f);  // punctual_light.h:36-39
        if (l.type == 2) {                                                 // punctual_light.h:28-34
            float ca = dot(normalize(-tl), f3(l.direction.x, l.direction.y, l.direction.z));
            float sp;
            if (ca < l.outer_cone) sp = 0.0f;
            else if (ca > l.inner_cone) sp = 1.0f;
            else {
                float t = (ca - l.outer_cone) / (l.inner_cone - l.outer_cone + 1e-7f);
                sp = t * t;
            }
            att *= sp;
        }
        inten = f3(l.color.x, l.color.y, l.color.z) * (l.intensity * att);
    } else {
        wi = -f3(l.direction.x, l.direction.y, l.direction.z);
        inten = f3(l.color.x, l.color.y, l.color.z) * l.intensity;
    }
    return dot(n, wi) > 0.0f;
}

// evaluatePunctualLights (direct_lighting.cuh:158-200)
template <bool COUNT, class Ctx>
__device__ __forceinline__ F3 punctual(const DevScene& s, F3 p, const MatEval<float>& m, F3 wo, F3 n, Ctx& c) {
    F3 ld = f3(0.0f);
    if (s.num_lights == 0 || s.lights == nullptr) return ld;
    for (uint32_t i = 0; i < s.num_lights; ++i) {
        F3 wi, inten;
        float dist;
        if (!punctual_light_sample(s, i, p, n, wi, inten, dist)) continue;
        if (occluded<COUNT>(s, make_shadow(p, n, wi, dist), c)) continue;
        BsdfEvalT<float> b = eval_bsdf(m, wo, wi, n, true);
        ld += inten * f3(b.f.x, b.f.y, b.f.z);
    }
    return ld;
}

// ---- surface interaction ----------------------------------------------------
struct Surface {
    F3 p, n;      // point, interpolated shading normal
    float u, v;   // texture coordinates
    float bu, bv; // barycentrics
    uint32_t tri; // BVH slot
    uint32_t mat;
    float t;
};

// Barycentric blend a w + b u + c v, every product and sum rounded on its own.
__device__ __forceinline__ float bary_rn(float a, float b, float c, float w, float u, float v) {
#pragma clang fp contract(off)
    return a * w + b * u + c * v;
}

// intersectTriangle's hit record (triangle.h:40-49).  The point, normal and
// texture coordinates round as the reference's host arithmetic does (no
// multiply-add fusion): a texture's filter weights are quantized to 1/256
// of a texel, and a last-bit difference in uv can move one step between
// neighbouring texels.
// The vertex data come from the triangle's shading record when the scene
// has them (one 64 B read after the hit), else through the face.
__device__ __forceinline__ Surface make_surface(const DevScene& s, F3 o, F3 d, const Hit& h) {
    Surface sf;
    float w;
    {
#pragma clang fp contract(off)
        w = 1.0f - h.u - h.v;
        sf.p = f3(o.x + h.t * d.x, o.y + h.t * d.y, o.z + h.t * d.z);
    }
    sf.t = h.t;
    F3 n0, n1, n2;
    float2 t0, t1, t2;
    if (s.shade_rec != nullptr) {
        const float4* r = s.shade_rec + 4 * static_cast<size_t>(h.tri);
        const float4 a = gld(r), b = gld(r + 1), c = gld(r + 2), e = gld(r + 3);
        n0 = f3(a.x, a.y, a.z);
        n1 = f3(a.w, b.x, b.y);
        n2 = f3(b.z, b.w, c.x);
        t0 = make_float2(c.y, c.z);
        t1 = make_float2(c.w, e.x);
        t2 = make_float2(e.y, e.z);
        sf.mat = __float_as_uint(e.w);
    } else {
        const uint4 f = s.faces[h.tri];
        n0 = f3(s.vnrm[f.x]);
        n1 = f3(s.vnrm[f.y]);
        n2 = f3(s.vnrm[f.z]);
        t0 = s.vuv[f.x];
        t1 = s.vuv[f.y];
        t2 = s.vuv[f.z];
        sf.mat = f.w;
    }
    sf.n = normalize_rn(f3(bary_rn(n0.x, n1.x, n2.x, w, h.u, h.v), bary_rn(n0.y, n1.y, n2.y, w, h.u, h.v),
                           bary_rn(n0.z, n1.z, n2.z, w, h.u, h.v)));
    sf.u = bary_rn(t0.x, t1.x, t2.x, w, h.u, h.v);
    sf.v = bary_rn(t0.y, t1.y, t2.y, w, h.u, h.v);
    sf.bu = h.u;
    sf.bv = h.v;
    sf.tri = h.tri;
    return sf;
}

// emissiveTrianglePdf (light_pdf.cuh:29-54)
__device__ __forceinline__ float emissive_pdf(const DevScene& s, uint32_t tri, float dist, F3 d) {
    if (tri >= s.num_faces || s.tri_pmf == nullptr) return 0.0f;
    float pmf = s.tri_pmf[tri];
    if (pmf <= 0.0f) return 0.0f;
    uint4 f = s.faces[tri];
    F3 p0 = f3(s.vpos[f.x]), p1 = f3(s.vpos[f.y]), p2 = f3(s.vpos[f.z]);
    F3 c = cross(p1 - p0, p2 - p0);
    float area = 0.5f * length(c);
    if (area <= 0.0f) return 0.0f;
    F3 ln = normalize(c);
    float cl = fabsf(dot(ln, -d));
    if (cl <= 1e-5f) return 0.0f;
    return pmf * dist * dist / (area * cl);
}

__device__ __forceinline__ bool mat_emissive(const pt_material& m) {
    return m.emissive_strength > 0.0f && (m.emissive.x > 0.0f || m.emissive.y > 0.0f || m.emissive.z > 0.0f);
}

// applyNormalMap (shading_utils.cuh:43-68); interpolateTangent (:29-41).
__device__ __forceinline__ F3 normal_map(const DevScene& s, const pt_material& m, F3 n, const Surface& sf, float u,
                                         float v) {
    if (!tex_valid(s, m.normal_tex_id)) return n;
    float4 nm = tex_fetch(s, m.normal_tex_id, u, v);
    F3 mn = f3(nm.x * 2.0f - 1.0f, nm.y * 2.0f - 1.0f, nm.z * 2.0f - 1.0f);
    mn.x *= m.normal_scale;
    mn.y *= m.normal_scale;
    mn = normalize(mn);
    F3 T = f3(0.0f);
    float sign = 1.0f;
    if (s.vtan != nullptr) {
        uint4 f = s.faces[sf.tri];
        float w = 1.0f - sf.bu - sf.bv;
        float4 a = s.vtan[f.x], b = s.vtan[f.y], c = s.vtan[f.z];
        T = normalize(f3(w * a.x + sf.bu * b.x + sf.bv * c.x, w * a.y + sf.bu * b.y + sf.bv * c.y,
                         w * a.z + sf.bu * b.z + sf.bv * c.z));
        sign = a.w >= 0.0f ? 1.0f : -1.0f;
    }
    if (!is_zero(T)) {
        T = normalize(T - n * dot(n, T));
        F3 B = cross(n, T) * sign;
        return normalize(T * mn.x + B * mn.y + n * mn.z);
    }
    Frame fr = make_frame(n);
    return normalize(to_world(fr, mn));
}

// ---- camera -------------------------------------------------------------------
// The thin lens of emitRay (path.cu:48-56): the origin moves to a point of
// the aperture disk, the ray passes through the focal point.
__device__ __forceinline__ void thin_lens(F3& pos, F3& dir, F3 fwd, F3 rgt, F3 cup, float focal_distance,
                                          float aperture, float l1, float l2) {
#pragma clang fp contract(off)
    const float k = focal_distance / dot_rn(dir, fwd);
    const F3 focal = f3(pos.x + dir.x * k, pos.y + dir.y * k, pos.z + dir.z * k);
    const float r = sqrtf(l1) * aperture;
    const float th = kTwoPi * l2;
    const float rc = r * cosf(th), rs = r * sinf(th);
    pos = f3(pos.x + (rgt.x * rc + cup.x * rs), pos.y + (rgt.y * rc + cup.y * rs), pos.z + (rgt.z * rc + cup.z * rs));
    dir = normalize_rn(f3(focal.x - pos.x, focal.y - pos.y, focal.z - pos.z));
}

// emitRay (path.cu:31-59), every product and sum rounded on its own as in
// the reference's host arithmetic (the ray decides which texels a lookup
// weights, to 1/256 of a texel).
__device__ __forceinline__ void camera_ray(const RenderParams& p, int x, int y, Sampler& rng, F3& o, F3& d) {
    float jx = rng.next();
    float jy = rng.next();
    F3 pos = f3(p.cam_pos[0], p.cam_pos[1], p.cam_pos[2]);
    const F3 fwd = f3(p.cam_fwd[0], p.cam_fwd[1], p.cam_fwd[2]);
    const F3 rgt = f3(p.cam_right[0], p.cam_right[1], p.cam_right[2]);
    const F3 cup = f3(p.cam_up[0], p.cam_up[1], p.cam_up[2]);
    F3 dir;
    {
#pragma clang fp contract(off)
        const float u = (2.0f * (static_cast<float>(x) + jx) / static_cast<float>(p.width) - 1.0f) * p.tan_half_fov *
                        p.aspect;
        const float v = (2.0f * (static_cast<float>(y) + jy) / static_cast<float>(p.height) - 1.0f) * p.tan_half_fov;
        dir = normalize_rn(f3(u * rgt.x + v * cup.x + fwd.x, u * rgt.y + v * cup.y + fwd.y,
                              u * rgt.z + v * cup.z + fwd.z));
    }
    if (p.aperture > 0.0f) {
        const float l1 = rng.next();
        const float l2 = rng.next();
        thin_lens(pos, dir, fwd, rgt, cup, p.focal_distance, p.aperture, l1, l2);
    }
    o = pos;
    d = normalize_rn(dir);  // Ray's constructor normalizes (core/ray.h:13)
}

// ---- path ---------------------------------------------------------------------
struct PathState {
    F3 o, d;      // current ray
    F3 tp;        // throughput
    F3 lo;        // radiance gathered so far
    float last_pdf;
    int depth;
    bool last_delta;
};

// What a bounce knows about its shading point, handed to the policy hooks.
struct ShadeCtx {
    const pt_material* mat;
    const MatEval<float>* me;
    F3 p, wo, n;   // point, outgoing direction, shading normal
    float tu, tv;  // texture coordinates after the uv transform
    bool entering;
};

// Visibility policy of the plain megakernel: shadow rays are traced inline
// and the contributions added at once (direct_lighting.cuh:220-249).
// A policy also receives hooks at the points where radiance or throughput
// changes (the adjoint uses them; here they are empty).
template <bool COUNT, class Ctx>
struct InlineNee {
    static constexpr bool kDrawAhead = true;
    Ctx& c;
    __device__ __forceinline__ explicit InlineNee(Ctx& ctx) : c(ctx) {}
    // The area term is added before the environment term, one at a time,
    // so the wavefront schedule (one shadow queue per term) sums alike.
    __device__ __forceinline__ void direct(const DevScene& s, PathState& st, bool va, const ShadowRay& sa, F3 ca,
                                           bool ve, const ShadowRay& se, F3 ce, const LightSample&,
                                           const LightSample&, const ShadeCtx&) {
        if (va && !occluded<COUNT>(s, sa, c)) st.lo += ca;
        if (ve && !occluded<COUNT>(s, se, c)) st.lo += ce;
    }
    __device__ __forceinline__ void punctual_lights(const DevScene& s, PathState& st, const ShadeCtx& x) {
        st.lo += st.tp * punctual<COUNT>(s, x.p, *x.me, x.wo, x.n, c);
    }
    __device__ __forceinline__ void on_env_miss(const DevScene&, const PathState&, float) {}
    __device__ __forceinline__ void on_bsdf_sample(const DevScene&, const PathState&, const BsdfSample&,
                                                   const ShadeCtx&) {}
};

// One bounce of pathTrace (path.cu:69-199) after the closest-hit query of the
// current ray: returns true when the path continues with the new ray in st.
// NEE contributions are premultiplied by the throughput and handed to the
// visibility policy together with their shadow rays.
// lo + (tp * e) * m with every product and sum rounded on its own: the
// passes that sum a path's radiance (the forward pass from a zero start plus
// the stored sum, the adjoint's replay from the stored sum) then round alike.
__device__ __forceinline__ F3 add_radiance(F3 lo, F3 tp, F3 e, float m) {
#pragma clang fp contract(off)
    return f3(lo.x + (tp.x * e.x) * m, lo.y + (tp.y * e.y) * m, lo.z + (tp.z * e.z) * m);
}

// A ray that escaped (path.cu:71-82): the environment's radiance, MIS-weighted
// against the environment sampling after a non-delta bounce; the path ends.
// In three steps, so that a kernel can keep the gathers of several escaped
// paths in flight at once: escape_taps (where to read: the bilinear taps of the
// direction's equirect coordinates and its pmf entry), the reads (four texels,
// and the pmf entry when env_pdf_wanted), env_escape (the radiance onto the
// path's sum).  escape() chains them for one path.
struct EscapeTaps {
    TexTaps tex;
    int pmf_idx;  // the direction's pmf entry, or -1: its pdf is 0 (no table, or a pole)
    float st;     // sin(theta) of the direction
};
__device__ __forceinline__ EscapeTaps escape_taps(const DevScene& s, F3 d) {
    EscapeTaps e;
    float u = 0.0f, v = 0.0f;
    dir_to_uv(d, u, v);
    e.tex = tex_taps(s.env_w, s.env_h, s.env_wrap_s, s.env_wrap_t, u, v);
    // environmentPdf (light_pdf.cuh:12-27), up to its table read
    e.pmf_idx = -1;
    e.st = 0.0f;
    if (s.env_pmf != nullptr && s.num_env_texels != 0 && s.env_w > 0 && s.env_h > 0) {
        int x = static_cast<int>(clampf(u * s.env_w, 0.0f, static_cast<float>(s.env_w - 1)));
        int y = static_cast<int>(clampf(v * s.env_h, 0.0f, static_cast<float>(s.env_h - 1)));
        e.st = sqrtf(fmaxf(0.0f, 1.0f - d.y * d.y));
        if (e.st > 1e-5f) e.pmf_idx = y * s.env_w + x;
    }
    return e;
}
// The MIS weight needs the environment pdf after a non-delta bounce only.
__device__ __forceinline__ bool env_pdf_wanted(const PathState& st) { return st.depth > 0 && !st.last_delta; }
template <class Nee>
__device__ __forceinline__ void env_escape(const DevScene& s, PathState& st, const EscapeTaps& e, const float4 (&t)[4],
                                           float pmf, Nee& nee) {
    const float4 c = texel_blend(e.tex.w[0], e.tex.w[1], e.tex.w[2], e.tex.w[3], t[0], t[1], t[2], t[3]);
    const F3 env = f3(c.x, c.y, c.z) * kTwoPi;
    if (!is_zero(env)) {
        float mis = 1.0f;
        if (env_pdf_wanted(st)) {
            const float pdf =
                e.pmf_idx < 0 ? 0.0f : pmf * static_cast<float>(s.env_w * s.env_h) / (kTwoPi * kPi * e.st);
            mis = power_heuristic(st.last_pdf, pdf);
        }
        nee.on_env_miss(s, st, mis);
        st.lo = add_radiance(st.lo, st.tp, env, mis);
    }
}
template <class Nee>
__device__ __forceinline__ void escape(const DevScene& s, PathState& st, Nee& nee) {
    if (s.env == nullptr) return;  // no environment: no radiance
    const EscapeTaps e = escape_taps(s, st.d);
    const float4 t[4] = {s.env[e.tex.idx[0]], s.env[e.tex.idx[1]], s.env[e.tex.idx[2]], s.env[e.tex.idx[3]]};
    const float pmf = env_pdf_wanted(st) && e.pmf_idx >= 0 ? s.env_pmf[e.pmf_idx] : 0.0f;
    env_escape(s, st, e, t, pmf, nee);
}

template <class Nee>
__device__ __forceinline__ bool bounce(const DevScene& s, const RenderParams& p, const Hit& h, PathState& st,
                                       Sampler& rng, Nee& nee) {
    const int depth = st.depth;
    if (h.tri == 0xFFFFFFFFu) {
        escape(s, st, nee);
        return false;
    }
    const bool has_area = s.num_emissive > 0 && s.emissive != nullptr;
    const bool has_env =
        s.env != nullptr && s.env_alias != nullptr && s.num_env_texels > 0 && s.env_w > 0 && s.env_h > 0;
    // Light samples drawn ahead from a copy of the sampler: their gathers
    // (alias table -> emitter or texel) overlap the surface's.  One is used
    // only where the path's own sampler reaches the same position (a blended
    // surface draws a number first), so it takes the same numbers through
    // the same code as drawing it in place, and the sampler's sequence is
    // unchanged.  Agreement to the last bit between two kernels still depends
    // on the compiler's multiply-add fusion, which follows the code's shape:
    // the megakernel and the wavefront shading kernel both draw ahead (their
    // images are identical); the adjoint replays draw in place, which keeps
    // them equal to the forward pass they replay.
    AreaGeom pre_a;
    EnvGeom pre_e;
    pre_a.key = pre_e.key = 0xFFFFFFFFu;
    bool ahead_a = false, ahead_e = false;
    if constexpr (Nee::kDrawAhead) {
        Sampler ahead = rng;
        if (has_area) {
            area_geom(s, ahead, pre_a);
            ahead_a = true;
        }
        if (has_env) {
            env_geom(s, ahead, pre_e);
            ahead_e = true;
        }
    }
    Surface sf = make_surface(s, st.o, st.d, h);
    const pt_material& mat = s.materials[sf.mat];
    bool entering = dot(sf.n, st.d) < 0.0f;
    F3 gn;
    if (mat.double_sided && !entering) {
        gn = -sf.n;
        entering = true;
    } else {
        gn = entering ? sf.n : -sf.n;
    }
    F3 wo = -st.d;
    float tu = sf.u, tv = sf.v;
    apply_uv_transform(mat, tu, tv);
    F3 n = normal_map(s, mat, gn, sf, tu, tv);
    MatEval<float> me = fetch_material(s, mat, tu, tv);
    if (mat_emissive(mat)) {
        float mis = 1.0f;
        if (depth > 0 && !st.last_delta) mis = power_heuristic(st.last_pdf, emissive_pdf(s, sf.tri, sf.t, st.d));
        st.lo = add_radiance(st.lo, st.tp, f3(mat.emissive.x, mat.emissive.y, mat.emissive.z) * mat.emissive_strength,
                             mis);
        return false;
    }
    if (mat.alpha_mode == 1) {  // path.cu:136-146 (raw uv)
        float alpha = 1.0f;
        if (tex_valid(s, mat.base_color_tex_id)) alpha = tex_fetch(s, mat.base_color_tex_id, sf.u, sf.v).w;
        if (alpha < mat.alpha_cutoff) {
            st.o = sf.p + st.d * 1e-4f;
            st.d = normalize(st.d);
            return true;
        }
    } else if (mat.alpha_mode == 2) {  // :147-158
        float alpha = mat.base_alpha;
        if (tex_valid(s, mat.base_color_tex_id)) alpha *= tex_fetch(s, mat.base_color_tex_id, sf.u, sf.v).w;
        if (rng.next() >= alpha) {
            st.o = sf.p + st.d * 1e-4f;
            st.d = normalize(st.d);
            return true;
        }
    }
    bool delta_surface =
        (me.metallic >= 1.0f && me.roughness < 0.01f) || (me.spec_trans >= 1.0f && me.roughness < 0.01f);
    ShadeCtx x;
    x.mat = &mat;
    x.me = &me;
    x.p = sf.p;
    x.wo = wo;
    x.n = n;
    x.tu = tu;
    x.tv = tv;
    x.entering = entering;
    if (!delta_surface) {  // :161-166
        ShadowRay sa, se;
        bool va = false, ve = false;
        F3 ca = f3(0.0f), ce = f3(0.0f);
        LightSample la, le;
        la.valid = le.valid = false;
        if (has_area) {
            if (ahead_a && rng.key() == pre_a.key) {
                rng.skip(4);
                la = area_finish(sf.p, n, pre_a, sa);
            } else {
                la = sample_area_nv(s, sf.p, n, rng, sa);
            }
            if (la.valid) {
                ca = st.tp * eval_direct(la, me, wo, n);
                va = !is_zero(ca);
            }
        }
        if (has_env) {
            if (ahead_e && rng.key() == pre_e.key) {
                rng.skip(4);
                le = env_finish(sf.p, n, pre_e, se);
            } else {
                le = sample_env_nv(s, sf.p, n, rng, se);
            }
            if (le.valid) {
                ce = st.tp * eval_direct(le, me, wo, n);
                ve = !is_zero(ce);
            }
        }
        nee.direct(s, st, va, sa, ca, ve, se, ce, la, le, x);
        if (s.num_lights > 0 && s.lights != nullptr) nee.punctual_lights(s, st, x);
    }
    BsdfSample bs = sample_bsdf(me, wo, n, rng, entering);
    if (!bs.valid) return false;
    nee.on_bsdf_sample(s, st, bs, x);
    st.tp *= bs.weight;
    bool transmitted = dot(bs.wi, gn) < 0.0f;
    F3 off = transmitted ? -gn : gn;
    st.o = sf.p + off * 1e-4f;
    st.d = normalize(bs.wi);
    if (transmitted && me.thickness > 0.0f && me.attenuation_distance < 1e20f) {  // :178-187
        float t = me.thickness, ad = me.attenuation_distance;
        F3 ac = me.attenuation_color;
        st.tp *= f3(expf(-t * fmaxf(-logf(fmaxf(ac.x, 1e-6f)), 0.0f) / ad),
                    expf(-t * fmaxf(-logf(fmaxf(ac.y, 1e-6f)), 0.0f) / ad),
                    expf(-t * fmaxf(-logf(fmaxf(ac.z, 1e-6f)), 0.0f) / ad));
    }
    st.last_pdf = bs.pdf;
    st.last_delta = bs.delta;
    if (depth > 2) {  // :193-198
        float surv = fminf(max3(st.tp), 0.95f);
        if (rng.next() >= surv) return false;
        st.tp /= surv;
    }
    return true;
}

__device__ __forceinline__ void begin_path(const RenderParams& p, int x, int y, Sampler& rng, PathState& st) {
    camera_ray(p, x, y, rng, st.o, st.d);
    st.tp = f3(1.0f);
    st.lo = f3(0.0f);
    st.last_pdf = 0.0f;
    st.last_delta = true;
    st.depth = 0;
}

// path.cu:201-204
__device__ __forceinline__ F3 clamp_radiance(F3 lo) {
    return f3(fminf(lo.x, 100.0f), fminf(lo.y, 100.0f), fminf(lo.z, 100.0f));
}

// One path with a given visibility policy; returns its radiance before the clamp.
template <bool COUNT, class Ctx, class Nee>
__device__ __forceinline__ F3 path_radiance(const DevScene& s, const RenderParams& p, int x, int y, Sampler& rng,
                                            Ctx& c, Nee& nee) {
    PathState st;
    begin_path(p, x, y, rng, st);
    for (; st.depth < p.max_depth; ++st.depth) {
        c.closest++;
        Hit h = trace_closest_v<Ctx::kWide, COUNT>(s, st.o, st.d, c.sk, c.cnt, c.overflow);
        if (!bounce(s, p, h, st, rng, nee)) break;
    }
    return st.lo;
}

template <bool COUNT, class Ctx>
__device__ F3 path_trace(const DevScene& s, const RenderParams& p, int x, int y, Sampler& rng, Ctx& c) {
    InlineNee<COUNT, Ctx> nee(c);
    return clamp_radiance(path_radiance<COUNT>(s, p, x, y, rng, c, nee));
}

// (the translation unit's default again)
#pragma clang fp contract(fast)

}  // namespace ptd

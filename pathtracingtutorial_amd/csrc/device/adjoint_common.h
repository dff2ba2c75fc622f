// Pieces of the path-replay adjoint shared by its megakernel and wavefront
// forms: the derivatives of one vertex's BSDF factor and the chain from the
// resolved material parameters to the parameter buffers.
#pragma once

#include "adjoint.h"
#include "dev_dual.h"
#include "dev_integrator.h"

namespace ptd {

// Bilinear environment taps of direction d (sampleEnvironmentRadiance's lookup).
__device__ __forceinline__ TexTaps env_taps(const DevScene& s, F3 d) {
    float u, v;
    dir_to_uv(d, u, v);
    return tex_taps(s.env_w, s.env_h, s.env_wrap_s, s.env_wrap_t, u, v);
}

// Sensitivities gv[] with respect to a vertex's resolved parameters (base
// colour rgb, roughness, metallic) mapped to the parameter buffers:
// material factors m[] and, for a base-colour texture, the weight k3 of the
// four bilinear texel taps (global texel index tidx[k], weight tw[k]).
struct ParamChain {
    float m[kDP];
    bool tex;
    uint32_t tidx[4];
    float tw[4];
    F3 k3;
};

__device__ __forceinline__ void param_chain(const DevScene& s, const pt_material& m, float tu, float tv,
                                            const float gv[kDP], ParamChain& o) {
    F3 T = f3(1.0f);
    const bool has_tex = s.num_textures > 0 && s.textures != nullptr && tex_valid(s, m.base_color_tex_id);
    float mr_y = 1.0f, mr_z = 1.0f;
    if (s.num_textures > 0 && s.textures != nullptr && tex_valid(s, m.metallic_roughness_tex_id)) {
        float4 mr = tex_fetch(s, m.metallic_roughness_tex_id, tu, tv);
        mr_y = mr.y;
        mr_z = mr.z;
    }
    o.tex = false;
    if (has_tex) {
        const TexDesc& td = s.textures[m.base_color_tex_id];
        TexTaps t = tex_taps(td.w, td.h, td.wrap_s, td.wrap_t, tu, tv);
        float4 raw = tex_fetch(s, m.base_color_tex_id, tu, tv);
        F3 dec = td.linear ? f3(1.0f)
                           : f3(2.2f * powf(raw.x, 1.2f), 2.2f * powf(raw.y, 1.2f), 2.2f * powf(raw.z, 1.2f));
        T = td.linear ? f3(raw.x, raw.y, raw.z) : f3(powf(raw.x, 2.2f), powf(raw.y, 2.2f), powf(raw.z, 2.2f));
        o.tex = true;
        o.k3 = f3(gv[0] * m.base_color.x * dec.x, gv[1] * m.base_color.y * dec.y, gv[2] * m.base_color.z * dec.z);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            o.tidx[k] = td.offset + t.idx[k];
            o.tw[k] = t.w[k];
        }
    }
    o.m[0] = gv[0] * T.x;
    o.m[1] = gv[1] * T.y;
    o.m[2] = gv[2] * T.z;
    const float rough_raw = m.roughness * mr_y, metal_raw = m.metallic * mr_z;
    // the clamps of fetchMaterial (disney_bsdf.cuh:118-120) stop the gradient outside their range
    o.m[3] = (rough_raw >= 0.001f && rough_raw <= 1.0f) ? gv[kDRough] * mr_y : 0.0f;
    o.m[4] = (metal_raw >= 0.0f && metal_raw <= 1.0f) ? gv[kDMetal] * mr_z : 0.0f;
}

// d/dtheta of  T * Le * f * mis / (pdf + 1e-7)  for a light sample, with the
// pixel adjoint g: gv[] += sum_c g_c T_c Le_c k df_c/dtheta.  Returns false
// when the BSDF pdf at the light direction is zero (no term).  env_coef =
// g * T * f * k, the weight of d/dLe for environment samples.
// (NP: the partials carried, see to_dual; gv[NP..kDP) are left alone.)
template <int NP = kDP>
__device__ __forceinline__ bool light_grad(const PathState& st, F3 g, const LightSample& ls, const ShadeCtx& x,
                                           float gv[kDP], F3& env_coef) {
    BsdfEvalT<Dual<NP>> bd = eval_bsdf(to_dual<NP>(*x.me), x.wo, ls.wi, x.n, true);
    const float pdf_b = bd.pdf.v;
    if (pdf_b <= 0.0f) return false;
    const float k = power_heuristic(ls.pdf, pdf_b) / (ls.pdf + 1e-7f);
    const F3 w = f3(g.x * st.tp.x * ls.radiance.x * k, g.y * st.tp.y * ls.radiance.y * k,
                    g.z * st.tp.z * ls.radiance.z * k);
#pragma unroll
    for (int j = 0; j < NP; ++j) gv[j] += w.x * bd.f.x.d[j] + w.y * bd.f.y.d[j] + w.z * bd.f.z.d[j];
    env_coef = f3(g.x * st.tp.x * bd.f.x.v * k, g.y * st.tp.y * bd.f.y.v * k, g.z * st.tp.z * bd.f.z.v * k);
    return true;
}

// d/dtheta of a punctual light's  T * I * f  (no MIS): gv[] += sum_c g_c T_c I_c df_c/dtheta.
template <int NP = kDP>
__device__ __forceinline__ void punctual_grad(const PathState& st, F3 g, F3 wi, F3 inten, const ShadeCtx& x,
                                              float gv[kDP]) {
    BsdfEvalT<Dual<NP>> bd = eval_bsdf(to_dual<NP>(*x.me), x.wo, wi, x.n, true);
    const F3 w = f3(g.x * st.tp.x * inten.x, g.y * st.tp.y * inten.y, g.z * st.tp.z * inten.z);
#pragma unroll
    for (int j = 0; j < NP; ++j) gv[j] += w.x * bd.f.x.d[j] + w.y * bd.f.y.d[j] + w.z * bd.f.z.d[j];
}

// f of the sampled continuation over dual numbers: the BSDF at bs.wi, or
// for a delta lobe the analytic weight (disney_bsdf.cuh:391-403).
template <int NP = kDP>
__device__ __forceinline__ SV3<Dual<NP>> continuation_f(const ShadeCtx& x, const BsdfSample& bs) {
    const MatEval<Dual<NP>> md = to_dual<NP>(*x.me);
    if (bs.delta) {
        if (dot(x.n, bs.wi) < 0.0f) return md.base;
        SV3<Dual<NP>> f0 = compute_f0(md);
        F3 H = normalize(x.wo + bs.wi);
        float t5 = schlick_t5(fmaxf(dot(H, bs.wi), 0.0f));
        return sv3(f0.x + (1.0f - f0.x) * t5, f0.y + (1.0f - f0.y) * t5, f0.z + (1.0f - f0.z) * t5);
    }
    return eval_bsdf(md, x.wo, bs.wi, x.n, x.entering).f;
}

}  // namespace ptd

// Path-replay backpropagation (PRB) of the integrator.
//
// For every path of a batch the kernel first runs the primal estimator (the
// megakernel's code and random numbers) to get the path radiance L, then
// replays the path from the same sampler state with an adjoint visibility
// policy that, at each vertex,
//   * differentiates the NEE and punctual terms  T * Le * f(theta) * w / p
//     (and T * Le(env texels) * f * w / p for environment samples),
//   * differentiates the environment seen on a miss  T * env(texels) * w,
//   * and, for the sampled continuation  T' = T * f(theta) / p, adds
//     (L - L_so_far) * (df/dtheta) / f : the radiance still to come times the
//     relative change of the throughput.
// Sampling pdfs, MIS weights and Russian-roulette probabilities are detached
// (their gradients are not taken), which keeps the estimator unbiased because
// the MIS weights sum to one.  The per-path clamp at 100 passes the gradient
// only for channels below it.  f's derivatives with respect to the resolved
// base colour, roughness and metallic come from eval_bsdf over dual numbers;
// the chain to material factors and base-colour texels is applied when a
// vertex's sensitivities are flushed.  Material sensitivities are summed in
// LDS per block, texel and environment gradients go to global memory with
// FP32 atomics (their order is not fixed: results vary in the last bits).
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "adjoint.h"
#include "dev_dual.h"
#include "dev_integrator.h"

namespace ptd {

namespace {

constexpr int kTile = 16;
constexpr int kThreads = kTile * kTile;
constexpr int kXcds = 8;
constexpr uint32_t kLdsMaterials = 64;  // material sensitivities summed in LDS up to this many materials

__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t nb) {
    uint32_t q = nb / kXcds, r = nb % kXcds;
    uint32_t k = b % kXcds, i = b / kXcds;
    return k * q + min(k, r) + i;
}

struct AdjPath {
    F3 L;  // primal radiance of the path (before the clamp)
    F3 g;  // d loss / d L per channel (zero where the clamp is active)
};

// Bilinear environment taps of direction d (sampleEnvironmentRadiance's lookup).
__device__ __forceinline__ TexTaps env_taps(const DevScene& s, F3 d) {
    float u, v;
    dir_to_uv(d, u, v);
    return tex_taps(s.env_w, s.env_h, s.env_wrap_s, s.env_wrap_t, u, v);
}

__device__ __forceinline__ void add_env_grad(const DevScene& s, const AdjointGrads& G, F3 d, F3 coef) {
    if (G.env == nullptr || s.env == nullptr) return;
    if (coef.x == 0.0f && coef.y == 0.0f && coef.z == 0.0f) return;
    TexTaps t = env_taps(s, d);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float w = t.w[k] * kTwoPi;  // radiance = bilinear texel * 2 pi
        if (w == 0.0f) continue;
        float* dst = G.env + 4 * static_cast<size_t>(t.idx[k]);
        atomicAdd(dst + 0, coef.x * w);
        atomicAdd(dst + 1, coef.y * w);
        atomicAdd(dst + 2, coef.z * w);
    }
}

template <bool COUNT>
struct AdjointNee {
    TraceCtx& c;
    const AdjPath& a;
    const AdjointGrads& G;
    float* macc;  // LDS material sensitivities [m][5], or null
    __device__ __forceinline__ AdjointNee(TraceCtx& ctx, const AdjPath& ap, const AdjointGrads& g, float* lds)
        : c(ctx), a(ap), G(g), macc(lds) {}

    // Applies the chain from resolved parameters to the parameter buffers.
    __device__ void flush(const DevScene& s, const ShadeCtx& x, const float gv[kDP]) {
        bool any = false;
#pragma unroll
        for (int k = 0; k < kDP; ++k) any = any || gv[k] != 0.0f;
        if (!any) return;
        const pt_material& m = *x.mat;
        const uint32_t mi = static_cast<uint32_t>(x.mat - s.materials);
        F3 T = f3(1.0f);
        const bool has_tex = s.num_textures > 0 && s.textures != nullptr && tex_valid(s, m.base_color_tex_id);
        float mr_y = 1.0f, mr_z = 1.0f;
        if (s.num_textures > 0 && s.textures != nullptr && tex_valid(s, m.metallic_roughness_tex_id)) {
            float4 mr = tex_fetch(s, m.metallic_roughness_tex_id, x.tu, x.tv);
            mr_y = mr.y;
            mr_z = mr.z;
        }
        if (has_tex) {
            const TexDesc& td = s.textures[m.base_color_tex_id];
            TexTaps t = tex_taps(td.w, td.h, td.wrap_s, td.wrap_t, x.tu, x.tv);
            float4 raw = tex_fetch(s, m.base_color_tex_id, x.tu, x.tv);
            F3 dec = td.linear ? f3(1.0f)
                               : f3(2.2f * powf(raw.x, 1.2f), 2.2f * powf(raw.y, 1.2f), 2.2f * powf(raw.z, 1.2f));
            T = td.linear ? f3(raw.x, raw.y, raw.z) : f3(powf(raw.x, 2.2f), powf(raw.y, 2.2f), powf(raw.z, 2.2f));
            if (G.tex != nullptr) {
                const F3 k3 = f3(gv[0] * m.base_color.x * dec.x, gv[1] * m.base_color.y * dec.y,
                                 gv[2] * m.base_color.z * dec.z);
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    if (t.w[k] == 0.0f) continue;
                    float* dst = G.tex + 4 * (static_cast<size_t>(td.offset) + t.idx[k]);
                    if (k3.x != 0.0f) atomicAdd(dst + 0, k3.x * t.w[k]);
                    if (k3.y != 0.0f) atomicAdd(dst + 1, k3.y * t.w[k]);
                    if (k3.z != 0.0f) atomicAdd(dst + 2, k3.z * t.w[k]);
                }
            }
        }
        float out[kDP];
        out[0] = gv[0] * T.x;
        out[1] = gv[1] * T.y;
        out[2] = gv[2] * T.z;
        const float rough_raw = m.roughness * mr_y, metal_raw = m.metallic * mr_z;
        // the clamps of fetchMaterial (disney_bsdf.cuh:118-120) stop the gradient outside their range
        out[3] = (rough_raw >= 0.001f && rough_raw <= 1.0f) ? gv[kDRough] * mr_y : 0.0f;
        out[4] = (metal_raw >= 0.0f && metal_raw <= 1.0f) ? gv[kDMetal] * mr_z : 0.0f;
        if (macc != nullptr && mi < kLdsMaterials) {
#pragma unroll
            for (int k = 0; k < kDP; ++k)
                if (out[k] != 0.0f) atomicAdd(&macc[mi * kDP + k], out[k]);
        } else {
            if (G.base) {
                for (int k = 0; k < 3; ++k)
                    if (out[k] != 0.0f) atomicAdd(&G.base[3 * mi + k], out[k]);
            }
            if (G.rough && out[3] != 0.0f) atomicAdd(&G.rough[mi], out[3]);
            if (G.metal && out[4] != 0.0f) atomicAdd(&G.metal[mi], out[4]);
        }
    }

    // d/dtheta of T * Le * f * mis / (pdf + 1e-7) at the light direction;
    // environment samples also pass d/dLe to the env texels.
    __device__ void light_grad(const DevScene& s, const PathState& st, const LightSample& ls, const ShadeCtx& x,
                               bool env, float gv[kDP]) {
        BsdfEvalT<D5> bd = eval_bsdf(to_dual(*x.me), x.wo, ls.wi, x.n, true);
        const float pdf_b = bd.pdf.v;
        if (pdf_b <= 0.0f) return;
        const float k = power_heuristic(ls.pdf, pdf_b) / (ls.pdf + 1e-7f);
        const F3 w = f3(a.g.x * st.tp.x * ls.radiance.x * k, a.g.y * st.tp.y * ls.radiance.y * k,
                        a.g.z * st.tp.z * ls.radiance.z * k);
#pragma unroll
        for (int j = 0; j < kDP; ++j) gv[j] += w.x * bd.f.x.d[j] + w.y * bd.f.y.d[j] + w.z * bd.f.z.d[j];
        if (env)
            add_env_grad(s, G, ls.wi,
                         f3(a.g.x * st.tp.x * bd.f.x.v * k, a.g.y * st.tp.y * bd.f.y.v * k,
                            a.g.z * st.tp.z * bd.f.z.v * k));
    }

    __device__ void direct(const DevScene& s, PathState& st, bool va, const ShadowRay& sa, F3 ca, bool ve,
                           const ShadowRay& se, F3 ce, const LightSample& la, const LightSample& le,
                           const ShadeCtx& x) {
        float gv[kDP] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        if (va && !occluded<COUNT>(s, sa, c)) {
            st.lo += ca;
            light_grad(s, st, la, x, false, gv);
        }
        if (ve && !occluded<COUNT>(s, se, c)) {
            st.lo += ce;
            light_grad(s, st, le, x, true, gv);
        }
        flush(s, x, gv);
    }

    // evaluatePunctualLights with the same arithmetic as the forward pass.
    __device__ void punctual_lights(const DevScene& s, PathState& st, const ShadeCtx& x) {
        float gv[kDP] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        F3 ld = f3(0.0f);
        for (uint32_t i = 0; i < s.num_lights; ++i) {
            F3 wi, inten;
            float dist;
            if (!punctual_light_sample(s, i, x.p, x.n, wi, inten, dist)) continue;
            if (occluded<COUNT>(s, make_shadow(x.p, x.n, wi, dist), c)) continue;
            BsdfEvalT<float> b = eval_bsdf(*x.me, x.wo, wi, x.n, true);
            ld += inten * f3(b.f.x, b.f.y, b.f.z);
            BsdfEvalT<D5> bd = eval_bsdf(to_dual(*x.me), x.wo, wi, x.n, true);
            const F3 w = f3(a.g.x * st.tp.x * inten.x, a.g.y * st.tp.y * inten.y, a.g.z * st.tp.z * inten.z);
#pragma unroll
            for (int j = 0; j < kDP; ++j) gv[j] += w.x * bd.f.x.d[j] + w.y * bd.f.y.d[j] + w.z * bd.f.z.d[j];
        }
        st.lo += st.tp * ld;
        flush(s, x, gv);
    }

    __device__ void on_env_miss(const DevScene& s, const PathState& st, float mis) {
        add_env_grad(s, G, st.d, f3(a.g.x * st.tp.x * mis, a.g.y * st.tp.y * mis, a.g.z * st.tp.z * mis));
    }

    // The continuation's throughput factor: (L - L_so_far) * (df/dtheta) / f.
    __device__ void on_bsdf_sample(const DevScene& s, const PathState& st, const BsdfSample& bs, const ShadeCtx& x) {
        const F3 coef = f3(a.g.x * (a.L.x - st.lo.x), a.g.y * (a.L.y - st.lo.y), a.g.z * (a.L.z - st.lo.z));
        if (coef.x == 0.0f && coef.y == 0.0f && coef.z == 0.0f) return;
        float gv[kDP] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        const MatEval<D5> md = to_dual(*x.me);
        SV3<D5> f;
        if (bs.delta) {  // disney_bsdf.cuh:391-403: the analytic weight
            if (dot(x.n, bs.wi) < 0.0f) {
                f = md.base;
            } else {
                SV3<D5> f0 = compute_f0(md);
                F3 H = normalize(x.wo + bs.wi);
                float t5 = schlick_t5(fmaxf(dot(H, bs.wi), 0.0f));
                f = sv3(f0.x + (1.0f - f0.x) * t5, f0.y + (1.0f - f0.y) * t5, f0.z + (1.0f - f0.z) * t5);
            }
        } else {
            f = eval_bsdf(md, x.wo, bs.wi, x.n, x.entering).f;
        }
        const float rx = f.x.v != 0.0f ? coef.x / f.x.v : 0.0f;
        const float ry = f.y.v != 0.0f ? coef.y / f.y.v : 0.0f;
        const float rz = f.z.v != 0.0f ? coef.z / f.z.v : 0.0f;
#pragma unroll
        for (int j = 0; j < kDP; ++j) gv[j] = rx * f.x.d[j] + ry * f.y.d[j] + rz * f.z.d[j];
        flush(s, x, gv);
    }
};

// A view of the batch: camera/batch parameters, its image gradient and
// (optionally) its replay image, and its first 16x16 tile in the grid.
struct AdjView {
    RenderParams p;
    const float4* img_grad;
    float4* replay;
    uint32_t tile_offset;
    uint32_t _pad;
};

template <bool COUNT>
__global__ __launch_bounds__(kThreads) void k_adjoint(DevScene s, const AdjView* __restrict__ views, uint32_t nviews,
                                                      AdjointGrads G, DevCounters* ctr) {
    __shared__ uint32_t stack[kStackSize * kThreads];
    __shared__ float macc[kLdsMaterials * kDP];
    const bool use_lds = s.num_materials <= kLdsMaterials;
    if (use_lds)
        for (uint32_t k = threadIdx.x; k < s.num_materials * kDP; k += blockDim.x) macc[k] = 0.0f;
    __syncthreads();
    uint32_t gtile = xcd_tile(blockIdx.x, gridDim.x);
    uint32_t vlo = 0, vhi = nviews;  // the view owning this tile
    while (vhi - vlo > 1) {
        uint32_t mid = (vlo + vhi) >> 1;
        if (views[mid].tile_offset <= gtile) vlo = mid;
        else vhi = mid;
    }
    const AdjView& view = views[vlo];
    const RenderParams& p = view.p;
    const float4* __restrict__ img_grad = view.img_grad;
    float4* __restrict__ replay = view.replay;
    const uint32_t tile = gtile - view.tile_offset;
    int tx = static_cast<int>(tile % static_cast<uint32_t>(p.tiles_x));
    int ty = static_cast<int>(tile / static_cast<uint32_t>(p.tiles_x));
    int x = tx * kTile + static_cast<int>(threadIdx.x & (kTile - 1));
    int y = ty * kTile + static_cast<int>(threadIdx.x / kTile);
    TraceCtx c;
    c.stk = stack + threadIdx.x;
    c.stride = kThreads;
    c.cnt.nodes = c.cnt.tris = 0;
    c.scnt.nodes = c.scnt.tris = 0;
    c.overflow = c.closest = c.shadow = 0;
    if (x < p.width && y < p.height) {
        const size_t idx = static_cast<size_t>(y) * p.width + x;
        const float4 gp = img_grad[idx];
        const float inv_spf = 1.0f / static_cast<float>(p.spf);
        const F3 gpix = f3(gp.x * inv_spf, gp.y * inv_spf, gp.z * inv_spf);
        F3 acc = f3(0.0f);
        for (int i = 0; i < p.spf; ++i) {
            Sampler rng;
            if (p.sampler == 1)
                rng.init_sobol(x, y, p.frame_count * static_cast<uint32_t>(p.spf) + static_cast<uint32_t>(i), p.seed);
            else
                rng.init_pcg(x, y, p.frame_count, static_cast<uint32_t>(i), p.seed);
            Sampler rng2 = rng;
            InlineNee<COUNT> fwd(c);
            AdjPath ap;
            ap.L = path_radiance<COUNT>(s, p, x, y, rng, c, fwd);
            // fminf(L, 100) passes the gradient below the clamp (path.cu:201-204)
            ap.g = f3(ap.L.x <= 100.0f ? gpix.x : 0.0f, ap.L.y <= 100.0f ? gpix.y : 0.0f,
                      ap.L.z <= 100.0f ? gpix.z : 0.0f);
            if (replay == nullptr && is_zero(ap.g)) continue;
            AdjointNee<COUNT> adj(c, ap, G, use_lds ? macc : nullptr);
            F3 Lr = path_radiance<COUNT>(s, p, x, y, rng2, c, adj);
            acc += clamp_radiance(Lr);
        }
        if (replay != nullptr) {
            acc /= static_cast<float>(p.spf);
            replay[idx] = make_float4(acc.x, acc.y, acc.z, 1.0f);
        }
    }
    __syncthreads();
    if (use_lds) {
        for (uint32_t k = threadIdx.x; k < s.num_materials * kDP; k += blockDim.x) {
            const float v = macc[k];
            if (v == 0.0f) continue;
            const uint32_t mi = k / kDP, j = k % kDP;
            if (j < 3) {
                if (G.base) atomicAdd(&G.base[3 * mi + j], v);
            } else if (j == kDRough) {
                if (G.rough) atomicAdd(&G.rough[mi], v);
            } else if (G.metal) {
                atomicAdd(&G.metal[mi], v);
            }
        }
    }
    if (COUNT) {
        unsigned long long v0 = wave_sum_u64(c.closest), v1 = wave_sum_u64(c.shadow),
                           v2 = wave_sum_u64(c.cnt.nodes), v3 = wave_sum_u64(c.cnt.tris),
                           v6 = wave_sum_u64(c.scnt.nodes), v7 = wave_sum_u64(c.scnt.tris),
                           v5 = wave_sum_u64(c.overflow);
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(&ctr->closest_rays, v0);
            atomicAdd(&ctr->shadow_rays, v1);
            atomicAdd(&ctr->node_visits, v2 + v6);
            atomicAdd(&ctr->tri_tests, v3 + v7);
            atomicAdd(&ctr->shadow_node_visits, v6);
            atomicAdd(&ctr->shadow_tri_tests, v7);
            atomicAdd(&ctr->stack_overflows, v5);
        }
    }
}

}  // namespace

hipError_t launch_adjoint(const DevScene& s, const std::vector<AdjointView>& in, const AdjointGrads& g,
                          DevCounters* ctr, bool count, hipStream_t stream) {
    if (in.empty()) return hipSuccess;
    std::vector<AdjView> views(in.size());
    uint32_t tiles = 0;
    for (size_t i = 0; i < in.size(); ++i) {
        AdjView& v = views[i];
        std::memset(&v, 0, sizeof(v));
        v.p = in[i].params;
        const uint32_t tx = static_cast<uint32_t>((v.p.width + kTile - 1) / kTile);
        const uint32_t ty = static_cast<uint32_t>((v.p.height + kTile - 1) / kTile);
        v.p.tiles_x = static_cast<int32_t>(tx);
        v.img_grad = in[i].img_grad;
        v.replay = in[i].replay;
        v.tile_offset = tiles;
        tiles += tx * ty;
    }
    AdjView* d_views = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&d_views), sizeof(AdjView) * views.size());
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(d_views, views.data(), sizeof(AdjView) * views.size(), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) {
        dim3 grid(tiles), block(kThreads);
        const uint32_t nv = static_cast<uint32_t>(views.size());
        if (count)
            hipLaunchKernelGGL(k_adjoint<true>, grid, block, 0, stream, s, d_views, nv, g, ctr);
        else
            hipLaunchKernelGGL(k_adjoint<false>, grid, block, 0, stream, s, d_views, nv, g, ctr);
        e = hipGetLastError();
    }
    const hipError_t es = hipStreamSynchronize(stream);
    (void)hipFree(d_views);
    return e != hipSuccess ? e : es;
}

}  // namespace ptd

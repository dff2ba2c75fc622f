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
#include "adjoint_common.h"

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
    static constexpr bool kDrawAhead = false;
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
        const uint32_t mi = static_cast<uint32_t>(x.mat - s.materials);
        ParamChain pc;
        param_chain(s, *x.mat, x.tu, x.tv, gv, pc);
        if (pc.tex && G.tex != nullptr) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (pc.tw[k] == 0.0f) continue;
                float* dst = G.tex + 4 * static_cast<size_t>(pc.tidx[k]);
                if (pc.k3.x != 0.0f) atomicAdd(dst + 0, pc.k3.x * pc.tw[k]);
                if (pc.k3.y != 0.0f) atomicAdd(dst + 1, pc.k3.y * pc.tw[k]);
                if (pc.k3.z != 0.0f) atomicAdd(dst + 2, pc.k3.z * pc.tw[k]);
            }
        }
        if (macc != nullptr && mi < kLdsMaterials) {
#pragma unroll
            for (int k = 0; k < kDP; ++k)
                if (pc.m[k] != 0.0f) atomicAdd(&macc[mi * kDP + k], pc.m[k]);
        } else {
            if (G.base) {
                for (int k = 0; k < 3; ++k)
                    if (pc.m[k] != 0.0f) atomicAdd(&G.base[3 * mi + k], pc.m[k]);
            }
            if (G.rough && pc.m[3] != 0.0f) atomicAdd(&G.rough[mi], pc.m[3]);
            if (G.metal && pc.m[4] != 0.0f) atomicAdd(&G.metal[mi], pc.m[4]);
        }
    }

    __device__ void direct(const DevScene& s, PathState& st, bool va, const ShadowRay& sa, F3 ca, bool ve,
                           const ShadowRay& se, F3 ce, const LightSample& la, const LightSample& le,
                           const ShadeCtx& x) {
        float gv[kDP] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        F3 ec;
        if (va && !occluded<COUNT>(s, sa, c)) {
            st.lo += ca;
            light_grad(st, a.g, la, x, gv, ec);
        }
        if (ve && !occluded<COUNT>(s, se, c)) {
            st.lo += ce;
            if (light_grad(st, a.g, le, x, gv, ec)) add_env_grad(s, G, le.wi, ec);
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
            punctual_grad(st, a.g, wi, inten, x, gv);
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
        const SV3<D5> f = continuation_f(x, bs);
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
    c.sk = TraceStack<kStackSize, kStackSize, uint32_t>{stack + threadIdx.x, kThreads, nullptr, 0};
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
            InlineNee<COUNT, TraceCtx> fwd(c);
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

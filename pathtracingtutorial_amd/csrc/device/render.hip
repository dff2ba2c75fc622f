// Plain megakernel integrator and single-stage probe kernels.
//
// k_render_plain replaces kernelPathTrace (Part_2/src/integrators/path.cu:
// 209-242): one thread per pixel runs spf whole paths, writes the batch mean
// to the "new" film and running-mean accumulates into "acc".  Blocks are
// 16x16 pixel tiles (four 64-wide waves of 16x4); the tile of a block is
// chosen so that each XCD (blocks are dispatched round-robin over the 8 XCDs)
// works on one contiguous band of the image and keeps its BVH working set in
// its own L2.  The traversal stack lives in LDS: 32 entries x 256 lanes.
#include <hip/hip_runtime.h>

#include "dev_integrator.h"
#include "render.h"

namespace ptd {

constexpr int kTile = 16;
constexpr int kThreads = kTile * kTile;
constexpr int kXcds = 8;

__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t nb) {
    uint32_t q = nb / kXcds, r = nb % kXcds;
    uint32_t k = b % kXcds, i = b / kXcds;
    return k * q + min(k, r) + i;
}

template <bool COUNT>
__device__ __forceinline__ void flush_counters(DevCounters* ctr, const TraceCtx& c, uint32_t paths) {
    if (!COUNT) return;
    unsigned long long v0 = wave_sum_u64(c.closest), v1 = wave_sum_u64(c.shadow), v2 = wave_sum_u64(c.cnt.nodes),
                       v3 = wave_sum_u64(c.cnt.tris), v4 = wave_sum_u64(paths), v5 = wave_sum_u64(c.overflow),
                       v6 = wave_sum_u64(c.scnt.nodes), v7 = wave_sum_u64(c.scnt.tris);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&ctr->closest_rays, v0);
        atomicAdd(&ctr->shadow_rays, v1);
        atomicAdd(&ctr->node_visits, v2 + v6);
        atomicAdd(&ctr->tri_tests, v3 + v7);
        atomicAdd(&ctr->paths, v4);
        atomicAdd(&ctr->stack_overflows, v5);
        atomicAdd(&ctr->shadow_node_visits, v6);
        atomicAdd(&ctr->shadow_tri_tests, v7);
    }
}

template <bool COUNT>
__global__ __launch_bounds__(kThreads) void k_render_plain(DevScene s, RenderParams p, float4* __restrict__ film_new,
                                                          float4* __restrict__ film_acc, DevCounters* ctr) {
    __shared__ uint32_t stack[kStackSize * kThreads];
    uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
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
    uint32_t paths = 0;
    if (x < p.width && y < p.height) {
        F3 acc = f3(0.0f);
        for (int i = 0; i < p.spf; ++i) {
            Sampler rng;
            if (p.sampler == 1)
                rng.init_sobol(x, y, p.frame_count * static_cast<uint32_t>(p.spf) + static_cast<uint32_t>(i), p.seed);
            else
                rng.init_pcg(x, y, p.frame_count, static_cast<uint32_t>(i), p.seed);
            acc += path_trace<COUNT>(s, p, x, y, rng, c);
        }
        paths = static_cast<uint32_t>(p.spf);
        acc /= static_cast<float>(p.spf);
        size_t idx = static_cast<size_t>(y) * p.width + x;
        float4 nw = make_float4(acc.x, acc.y, acc.z, 1.0f);
        film_new[idx] = nw;
        float4 a = film_acc[idx];
        float inv = 1.0f / static_cast<float>(p.frame_count + 1u);
        a.x = a.x * (1.0f - inv) + nw.x * inv;
        a.y = a.y * (1.0f - inv) + nw.y * inv;
        a.z = a.z * (1.0f - inv) + nw.z * inv;
        a.w = a.w * (1.0f - inv) + nw.w * inv;
        film_acc[idx] = a;
    }
    flush_counters<COUNT>(ctr, c, paths);
}

hipError_t launch_render_plain(const DevScene& s, const RenderParams& p, float4* film_new, float4* film_acc,
                               DevCounters* ctr, bool count, hipStream_t stream) {
    uint32_t tiles_x = static_cast<uint32_t>((p.width + kTile - 1) / kTile);
    uint32_t tiles_y = static_cast<uint32_t>((p.height + kTile - 1) / kTile);
    RenderParams q = p;
    q.tiles_x = static_cast<int32_t>(tiles_x);
    dim3 grid(tiles_x * tiles_y), block(kThreads);
    if (count)
        hipLaunchKernelGGL(k_render_plain<true>, grid, block, 0, stream, s, q, film_new, film_acc, ctr);
    else
        hipLaunchKernelGGL(k_render_plain<false>, grid, block, 0, stream, s, q, film_new, film_acc, ctr);
    return hipGetLastError();
}

// ---- probes ---------------------------------------------------------------------

__global__ void k_probe_sampler(int type, uint32_t seed, const int32_t* __restrict__ xyi, int n, int dims,
                                uint32_t* __restrict__ out_u, float* __restrict__ out_f) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int x = xyi[3 * i], y = xyi[3 * i + 1];
    uint32_t idx = static_cast<uint32_t>(xyi[3 * i + 2]);
    Sampler a, b;
    if (type == 1) {
        a.init_sobol(x, y, idx, seed);
        b.init_sobol(x, y, idx, seed);
    } else {
        a.init_pcg(x, y, idx >> 16, idx & 0xFFFFu, seed);
        b.init_pcg(x, y, idx >> 16, idx & 0xFFFFu, seed);
    }
    for (int k = 0; k < dims; ++k) {
        out_u[static_cast<size_t>(i) * dims + k] = a.next_u32();
        out_f[static_cast<size_t>(i) * dims + k] = b.next();
    }
}

// One ray per thread; WIDE traverses the 4-wide BVH (its whole stack in LDS).
template <bool WIDE, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_trace_rays(DevScene s, const float* __restrict__ rays8, int n, int any,
                                                      float* __restrict__ t_out, int32_t* __restrict__ face_out,
                                                      float* __restrict__ bary_out) {
    constexpr int kTot = WIDE ? kWideStack : kStackSize;
    __shared__ uint32_t stack[kTot * BLOCK];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* r = rays8 + 8 * static_cast<size_t>(i);
    F3 o = f3(r[0], r[1], r[2]);
    F3 d = normalize_rn(f3(r[3], r[4], r[5]));  // Ray's constructor (core/ray.h:13)
    TraceCounters cnt{0, 0};
    uint32_t ovf = 0;
    const TraceStack<kTot, kTot> sk{stack + threadIdx.x, BLOCK, nullptr, 0};
    if (any) {
        bool occ = trace_any_v<WIDE, false>(s, o, d, r[6], sk, cnt, ovf);
        t_out[i] = occ ? 1.0f : 0.0f;
        face_out[i] = -1;
        return;
    }
    Hit h = trace_closest_v<WIDE, false>(s, o, d, sk, cnt, ovf);
    t_out[i] = h.t;
    if (h.tri == 0xFFFFFFFFu) {
        face_out[i] = -1;
    } else {
        face_out[i] = static_cast<int32_t>(__float_as_uint(s.tris[3 * h.tri].w) & 0x7FFFFFFFu);
    }
    if (bary_out) {
        bary_out[2 * i] = h.u;
        bary_out[2 * i + 1] = h.v;
    }
}

// Disney BSDF of one material (fetch + clamps, no textures): eval at (wo, wi)
// and a sample drawn with three given uniform numbers (choice, then the two
// warp numbers), as sampleDisneyBsdf<FixedRng> consumes them.
struct FixedRng {
    const float* u;
    int k;
    __device__ __forceinline__ float next() { return u[k++]; }
};

__global__ void k_probe_bsdf(const pt_material* __restrict__ mats, int n, const float* __restrict__ dirs,
                             const int32_t* __restrict__ entering, const float* __restrict__ us,
                             float* __restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    DevScene empty;
    memset(&empty, 0, sizeof(empty));
    MatEval<float> me = fetch_material(empty, mats[i], 0.0f, 0.0f);
    const float* d = dirs + 9 * static_cast<size_t>(i);
    F3 wo = f3(d[0], d[1], d[2]), wi = f3(d[3], d[4], d[5]), nn = f3(d[6], d[7], d[8]);
    const bool ent = entering[i] != 0;
    BsdfEvalT<float> e = eval_bsdf(me, wo, wi, nn, ent);
    FixedRng rng{us + 3 * static_cast<size_t>(i), 0};
    BsdfSample s = sample_bsdf(me, wo, nn, rng, ent);
    float* o = out + 12 * static_cast<size_t>(i);
    o[0] = e.f.x; o[1] = e.f.y; o[2] = e.f.z; o[3] = e.pdf;
    o[4] = s.wi.x; o[5] = s.wi.y; o[6] = s.wi.z;
    o[7] = s.weight.x; o[8] = s.weight.y; o[9] = s.weight.z;
    o[10] = s.pdf;
    o[11] = static_cast<float>((s.valid ? 1 : 0) | (s.delta ? 2 : 0));
}

hipError_t launch_probe_bsdf(const pt_material* d_mats, int n, const float* d_dirs, const int32_t* d_entering,
                             const float* d_u, float* d_out, hipStream_t stream) {
    hipLaunchKernelGGL(k_probe_bsdf, dim3((n + 127) / 128), dim3(128), 0, stream, d_mats, n, d_dirs, d_entering, d_u,
                       d_out);
    return hipGetLastError();
}

hipError_t launch_probe_sampler(int type, uint32_t seed, const int32_t* d_xyi, int n, int dims, uint32_t* d_u,
                                float* d_f, hipStream_t stream) {
    hipLaunchKernelGGL(k_probe_sampler, dim3((n + 255) / 256), dim3(256), 0, stream, type, seed, d_xyi, n, dims, d_u,
                       d_f);
    return hipGetLastError();
}

hipError_t launch_trace_rays(const DevScene& s, const float* d_rays8, int n, int any, bool wide, float* d_t,
                             int32_t* d_face, float* d_bary, hipStream_t stream) {
    if (wide)
        hipLaunchKernelGGL((k_trace_rays<true, 128>), dim3((n + 127) / 128), dim3(128), 0, stream, s, d_rays8, n, any,
                           d_t, d_face, d_bary);
    else
        hipLaunchKernelGGL((k_trace_rays<false, 256>), dim3((n + 255) / 256), dim3(256), 0, stream, s, d_rays8, n, any,
                           d_t, d_face, d_bary);
    return hipGetLastError();
}

// ---- traversal layout of the triangles -----------------------------------------

__global__ void k_emit_tris(const float4* __restrict__ vpos, const uint4* __restrict__ faces,
                            const uint32_t* __restrict__ order, uint32_t n, float4* __restrict__ tris,
                            uint4* __restrict__ faces_bvh) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t orig = order[i];
    uint4 f = faces[orig];
    float4 a = vpos[f.x], b = vpos[f.y], c = vpos[f.z];
    tris[3 * i + 0] = make_float4(a.x, a.y, a.z, __uint_as_float(orig));
    tris[3 * i + 1] = make_float4(b.x - a.x, b.y - a.y, b.z - a.z, 0.0f);
    tris[3 * i + 2] = make_float4(c.x - a.x, c.y - a.y, c.z - a.z, 0.0f);
    faces_bvh[i] = f;
}

__global__ void k_mark_leaf_ends(const float4* __restrict__ nodes, uint32_t m, float4* __restrict__ tris) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m || k == 1) return;
    uint32_t d0 = __float_as_uint(nodes[2 * k].w);
    if (!(d0 & kLeafFlag)) return;
    uint32_t first = d0 & ~kLeafFlag, count = __float_as_uint(nodes[2 * k + 1].w);
    if (count == 0) return;
    float4* t = tris + 3 * (first + count - 1);
    t->w = __uint_as_float(__float_as_uint(t->w) | 0x80000000u);
}

hipError_t launch_emit_tris(const float4* d_vpos, const uint4* d_faces_orig, const uint32_t* d_prim_order,
                            const float4* d_nodes, uint32_t num_nodes, uint32_t nfaces, float4* d_tris,
                            uint4* d_faces_bvh, hipStream_t stream) {
    if (nfaces == 0) return hipSuccess;
    hipLaunchKernelGGL(k_emit_tris, dim3((nfaces + 255) / 256), dim3(256), 0, stream, d_vpos, d_faces_orig,
                       d_prim_order, nfaces, d_tris, d_faces_bvh);
    hipLaunchKernelGGL(k_mark_leaf_ends, dim3((num_nodes + 255) / 256), dim3(256), 0, stream, d_nodes, num_nodes,
                       d_tris);
    return hipGetLastError();
}

// ---- copy probe ------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_copy_probe(const float4* __restrict__ src, float4* __restrict__ dst,
                                                    size_t n) {
    size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

hipError_t launch_copy_probe(const float4* src, float4* dst, size_t n, hipStream_t stream) {
    int blocks = 256 * 8 * 4;  // 4 blocks per CU per wave slot group, grid-stride over the buffer
    hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, stream, src, dst, n);
    return hipGetLastError();
}

}  // namespace ptd

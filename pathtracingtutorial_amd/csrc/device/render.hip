// Plain megakernel integrator and single-stage probe kernels.
//
// k_render_plain replaces kernelPathTrace (Part_2/src/integrators/path.cu:
// 209-242): one thread per pixel runs spf whole paths, writes the batch mean
// to the "new" film and running-mean accumulates into "acc".  The grid is
// persistent (as many 256-thread blocks as fit on the GPU at once); a block
// renders 16x16 pixel tiles (four 64-wide waves of 16x4) taken one at a time
// from a queue.  The tiles form 8 contiguous bands of the image, one per XCD
// (blocks are dispatched round-robin over the 8 XCDs, so block b starts on
// band b % 8 and keeps its BVH working set in its XCD's L2); a block whose
// band is exhausted takes tiles of the other bands, so uneven tiles balance.
// The traversal walks the 4-wide BVH (or the binary one when the scene asks
// for it) with a per-thread stack of (node, entry distance) entries: the
// first kMegaLds in LDS, the rest in device memory.
#include <hip/hip_runtime.h>

#include "dev_integrator.h"
#include "render.h"

namespace ptd {

constexpr int kTile = 16;
constexpr int kThreads = kTile * kTile;
constexpr int kXcds = 8;
constexpr int kMegaLds = 16;

template <bool WIDE>
constexpr int mega_stack_total() {
    return WIDE ? kWideStack : kStackSize;
}

template <bool COUNT, class Ctx>
__device__ __forceinline__ void flush_counters(DevCounters* ctr, const Ctx& c, uint32_t paths) {
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

__global__ void k_clear_u32(uint32_t* p, int n) {
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) p[i] = 0u;
}

// At 4 waves per SIMD (128 VGPRs, the rest of the path state spilled to
// scratch) the latency of the gathers hides better than at the 2 waves its
// ~230 registers would allow (537 -> 385 ms per 1080p 64-spp frame of the
// 1.3M-triangle sphere scene).
template <bool COUNT, bool WIDE>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_render_plain(DevScene s, RenderParams p, float4* __restrict__ film_new,
                                                          float4* __restrict__ film_acc, DevCounters* ctr,
                                                          uint32_t* __restrict__ tile_ctr, DistEntry* spill) {
    __shared__ DistEntry stack[kMegaLds * kThreads];
    __shared__ uint32_t grabbed;
    using SK = TraceStack<kMegaLds, mega_stack_total<WIDE>(), DistEntry>;
    TraceCtxT<WIDE, SK> c;
    c.sk = SK{stack + threadIdx.x, kThreads, spill + blockIdx.x * kThreads + threadIdx.x, gridDim.x * kThreads};
    c.cnt.nodes = c.cnt.tris = 0;
    c.scnt.nodes = c.scnt.tris = 0;
    c.overflow = c.closest = c.shadow = 0;
    uint32_t paths = 0;
    const uint32_t tiles_x = static_cast<uint32_t>(p.tiles_x);
    const uint32_t ntiles = tiles_x * static_cast<uint32_t>((p.height + kTile - 1) / kTile);
    const uint32_t q = ntiles / kXcds, r = ntiles % kXcds;
    const uint32_t home = blockIdx.x % kXcds;
    for (uint32_t k = 0; k < kXcds;) {
        const uint32_t band = (home + k) % kXcds;
        if (threadIdx.x == 0) grabbed = atomicAdd(&tile_ctr[band], 1u);
        __syncthreads();
        const uint32_t t = grabbed;  // the same for the whole block
        __syncthreads();
        if (t >= q + (band < r ? 1u : 0u)) {  // band exhausted: the next one
            ++k;
            continue;
        }
        const uint32_t tile = band * q + min(band, r) + t;
        const int x = static_cast<int>((tile % tiles_x) * kTile + (threadIdx.x & (kTile - 1)));
        const int y = static_cast<int>((tile / tiles_x) * kTile + threadIdx.x / kTile);
        if (x < p.width && y < p.height) {
            F3 acc = f3(0.0f);
            for (int i = 0; i < p.spf; ++i) {
                Sampler rng;
                if (p.sampler == 1)
                    rng.init_sobol(x, y, p.frame_count * static_cast<uint32_t>(p.spf) + static_cast<uint32_t>(i),
                                   p.seed);
                else
                    rng.init_pcg(x, y, p.frame_count, static_cast<uint32_t>(i), p.seed);
                acc += path_trace<COUNT>(s, p, x, y, rng, c);
            }
            paths += static_cast<uint32_t>(p.spf);
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
    }
    flush_counters<COUNT>(ctr, c, paths);
}

namespace {

template <bool COUNT, bool WIDE>
const void* plain_kernel() {
    return reinterpret_cast<const void*>(&k_render_plain<COUNT, WIDE>);
}

hipError_t plain_blocks(bool wide, int* blocks) {
    int dev = 0, cus = 0, per_cu = 0, per_cu_c = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    // the grid serves the counting and the plain kernel: the smaller occupancy of the two
    if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, wide ? plain_kernel<false, true>()
                                                                       : plain_kernel<false, false>(),
                                                         kThreads, 0);
    if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_c, wide ? plain_kernel<true, true>()
                                                                         : plain_kernel<true, false>(),
                                                         kThreads, 0);
    if (e != hipSuccess) return e;
    const int n = std::max(1, std::min(per_cu, per_cu_c)) * std::max(1, cus);
    *blocks = (n + kXcds - 1) / kXcds * kXcds;
    return hipSuccess;
}

}  // namespace

hipError_t render_plain_scratch(bool wide, size_t* spill_entries, int* tile_counters) {
    int blocks = 0;
    hipError_t e = plain_blocks(wide, &blocks);
    if (e != hipSuccess) return e;
    const int total = wide ? kWideStack : kStackSize;
    *spill_entries = static_cast<size_t>(blocks) * kThreads * static_cast<size_t>(total - kMegaLds);
    *tile_counters = kXcds;
    return hipSuccess;
}

hipError_t launch_render_plain(const DevScene& s, const RenderParams& p, float4* film_new, float4* film_acc,
                               DevCounters* ctr, bool count, bool wide, const PlainScratch& scratch,
                               hipStream_t stream) {
    if (wide && (s.wnodes == nullptr || s.num_wnodes == 0)) wide = false;
    int blocks = 0;
    hipError_t e = plain_blocks(wide, &blocks);
    if (e != hipSuccess) return e;
    const int total = wide ? kWideStack : kStackSize;
    const size_t need = static_cast<size_t>(blocks) * kThreads * static_cast<size_t>(total - kMegaLds);
    if (scratch.tile_ctr == nullptr || scratch.spill == nullptr || scratch.spill_entries < need)
        return hipErrorInvalidValue;
    RenderParams q = p;
    q.tiles_x = (p.width + kTile - 1) / kTile;
    hipLaunchKernelGGL(k_clear_u32, dim3(1), dim3(64), 0, stream, scratch.tile_ctr, kXcds);
    const dim3 grid(static_cast<uint32_t>(blocks)), block(kThreads);
    if (count && wide)
        hipLaunchKernelGGL((k_render_plain<true, true>), grid, block, 0, stream, s, q, film_new, film_acc, ctr,
                           scratch.tile_ctr, scratch.spill);
    else if (count)
        hipLaunchKernelGGL((k_render_plain<true, false>), grid, block, 0, stream, s, q, film_new, film_acc, ctr,
                           scratch.tile_ctr, scratch.spill);
    else if (wide)
        hipLaunchKernelGGL((k_render_plain<false, true>), grid, block, 0, stream, s, q, film_new, film_acc, ctr,
                           scratch.tile_ctr, scratch.spill);
    else
        hipLaunchKernelGGL((k_render_plain<false, false>), grid, block, 0, stream, s, q, film_new, film_acc, ctr,
                           scratch.tile_ctr, scratch.spill);
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
    __shared__ DistEntry stack[kTot * BLOCK];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* r = rays8 + 8 * static_cast<size_t>(i);
    F3 o = f3(r[0], r[1], r[2]);
    F3 d = normalize_rn(f3(r[3], r[4], r[5]));  // Ray's constructor (core/ray.h:13)
    TraceCounters cnt{0, 0};
    uint32_t ovf = 0;
    const TraceStack<kTot, kTot, DistEntry> sk{stack + threadIdx.x, BLOCK, nullptr, 0};
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

// ---- shading records ---------------------------------------------------------------
// DevScene::shade_rec from the BVH-ordered faces and the vertex normals and
// uvs (the same values make_surface would gather through the face).

__global__ __launch_bounds__(256) void k_shade_records(const uint4* __restrict__ faces,
                                                       const float4* __restrict__ vnrm, const float2* __restrict__ vuv,
                                                       uint32_t n, float4* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint4 f = faces[i];
    const float4 n0 = vnrm[f.x], n1 = vnrm[f.y], n2 = vnrm[f.z];
    const float2 t0 = vuv[f.x], t1 = vuv[f.y], t2 = vuv[f.z];
    float4* r = out + 4 * static_cast<size_t>(i);
    r[0] = make_float4(n0.x, n0.y, n0.z, n1.x);
    r[1] = make_float4(n1.y, n1.z, n2.x, n2.y);
    r[2] = make_float4(n2.z, t0.x, t0.y, t1.x);
    r[3] = make_float4(t1.y, t2.x, t2.y, __uint_as_float(f.w));
}

hipError_t launch_shade_records(const uint4* d_faces_bvh, const float4* d_vnrm, const float2* d_vuv, uint32_t nfaces,
                                float4* d_out, hipStream_t stream) {
    if (nfaces == 0) return hipSuccess;
    hipLaunchKernelGGL(k_shade_records, dim3((nfaces + 255) / 256), dim3(256), 0, stream, d_faces_bvh, d_vnrm, d_vuv,
                       nfaces, d_out);
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

// ---- reciprocal probe ------------------------------------------------------------
// Thread i takes significand i and walks every biased exponent 1..252 of both
// signs (|x| in [2^-126, 2^126)), comparing recip_rn(x) with IEEE 1.0f / x
// bit for bit.  The grid covers exactly 2^23 threads.

__global__ __launch_bounds__(256) void k_probe_reciprocal(unsigned long long* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (1u << 23)) return;
    unsigned long long bad = 0;
    for (uint32_t e = 1; e <= 252; ++e) {
        for (uint32_t s = 0; s < 2; ++s) {
            const float x = __uint_as_float((s << 31) | (e << 23) | i);
            bad += __float_as_uint(recip_rn(x)) != __float_as_uint(1.0f / x);
        }
    }
    if (bad) atomicAdd(out, bad);
}

hipError_t launch_probe_reciprocal(unsigned long long* d_out, hipStream_t stream) {
    hipLaunchKernelGGL(k_probe_reciprocal, dim3((1u << 23) / 256), dim3(256), 0, stream, d_out);
    return hipGetLastError();
}

}  // namespace ptd

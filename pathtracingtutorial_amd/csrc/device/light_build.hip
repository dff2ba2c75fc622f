// Parallel alias-table construction (see light_build.h).
//
// With p_i = w_i * n / W, bucket i is "light" when p_i < 1 (it needs 1 - p_i
// from another item) and "heavy" otherwise (it has p_i - 1 to give).  Laying
// the light deficits end to end in index order gives a line [0, T); the heavy
// surpluses, also in index order, tile the same line.  Light l whose deficit
// starts at s_l takes its alias from the heavy h whose surplus interval
// [S_h, S_{h+1}) contains s_l.  A light interval that runs past S_{h+1}
// overdraws heavy h by e - S_{h+1} (e = the light interval's end); heavy h's
// own bucket then keeps 1 - (e - S_{h+1}) of itself and takes the rest from
// heavy h+1, whose surplus interval starts exactly where that overdraw
// begins.  Every item ends with total mass p_i across the buckets, so the
// table samples pmf exactly (up to the rounding of the sums, carried in
// double precision: this is a table build, not a render kernel).
#include "light_build.h"

#include <mutex>
#include <vector>

namespace ptd {

namespace {

constexpr int kBlock = 256;
constexpr int kPerThread = 8;
constexpr int kTile = kBlock * kPerThread;

struct Acc {
    uint32_t lights;  // light items
    double def;       // light deficits
    double sur;       // heavy surpluses
};
__device__ __forceinline__ Acc acc_add(const Acc& a, const Acc& b) { return Acc{a.lights + b.lights, a.def + b.def, a.sur + b.sur}; }

// Emissive faces: how many, and their weight.
struct Cnt {
    uint32_t n;
    double w;
};
__device__ __forceinline__ Cnt acc_add(const Cnt& a, const Cnt& b) { return Cnt{a.n + b.n, a.w + b.w}; }

__device__ __forceinline__ Acc item_acc(const double* p, uint32_t i, uint32_t n) {
    if (i >= n) return Acc{0u, 0.0, 0.0};
    const double v = p[i];
    return v < 1.0 ? Acc{1u, 1.0 - v, 0.0} : Acc{0u, 0.0, v - 1.0};
}

// Inclusive scan of one value per thread over the block (Hillis-Steele in LDS).
template <class T>
__device__ T block_scan(T v, T* sh) {
    const int t = static_cast<int>(threadIdx.x);
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
        T add = t >= off ? sh[t - off] : T{};
        __syncthreads();
        if (t >= off) sh[t] = acc_add(sh[t], add);
        __syncthreads();
    }
    const T r = sh[t];
    __syncthreads();
    return r;
}

__device__ double block_sum(double v, double* sh) {
    const int t = static_cast<int>(threadIdx.x);
    sh[t] = v;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] += sh[t + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kBlock) void k_env_weights(const float4* __restrict__ env, int w, int h, float* wts) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t n = static_cast<uint32_t>(w) * static_cast<uint32_t>(h);
    if (i >= n) return;
    const int y = static_cast<int>(i / static_cast<uint32_t>(w));
    const float theta = 3.14159265358979323846f * (static_cast<float>(y) + 0.5f) / static_cast<float>(h);
    const float st = fmaxf(sinf(theta), 0.0f);
    const float4 c = env[i];
    wts[i] = fmaxf(0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z, 0.0f) * st;
}

// Partial sums of the weights per tile.
__global__ __launch_bounds__(kBlock) void k_sum_tiles(const float* __restrict__ w, uint32_t n, double* part) {
    __shared__ double sh[kBlock];
    const uint32_t base = blockIdx.x * kTile + threadIdx.x;
    double v = 0.0;
    for (int k = 0; k < kPerThread; ++k) {
        const uint32_t i = base + k * kBlock;
        if (i < n) v += static_cast<double>(fmaxf(w[i], 0.0f));
    }
    v = block_sum(v, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ __launch_bounds__(kBlock) void k_sum_final(const double* part, uint32_t np, double* total) {
    __shared__ double sh[kBlock];
    double v = 0.0;
    for (uint32_t i = threadIdx.x; i < np; i += kBlock) v += part[i];
    v = block_sum(v, sh);
    if (threadIdx.x == 0) *total = v;
}

__global__ __launch_bounds__(kBlock) void k_scale(const float* __restrict__ w, uint32_t n, const double* total,
                                                  double* p, float* pmf) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double W = *total;
    const double wi = static_cast<double>(fmaxf(w[i], 0.0f));
    p[i] = wi * static_cast<double>(n) / W;
    pmf[i] = static_cast<float>(wi / W);
}

// Scan phase 1: per-tile totals of (lights, deficits, surpluses).
__global__ __launch_bounds__(kBlock) void k_scan_tiles(const double* __restrict__ p, uint32_t n, Acc* part) {
    __shared__ Acc sh[kBlock];
    const uint32_t base = blockIdx.x * kTile + threadIdx.x * kPerThread;
    Acc v{0u, 0.0, 0.0};
    for (int k = 0; k < kPerThread; ++k) v = acc_add(v, item_acc(p, base + k, n));
    v = block_scan(v, sh);
    if (threadIdx.x == kBlock - 1) part[blockIdx.x] = v;
}

// Scan phase 2: exclusive offsets of the tiles (one block); the grand total in part[np].
__global__ __launch_bounds__(kBlock) void k_scan_parts(Acc* part, uint32_t np) {
    __shared__ Acc sh[kBlock];
    Acc carry{0u, 0.0, 0.0};
    for (uint32_t b0 = 0; b0 < np; b0 += kBlock) {
        const uint32_t i = b0 + threadIdx.x;
        const Acc v = i < np ? part[i] : Acc{0u, 0.0, 0.0};
        const Acc inc = block_scan(v, sh);
        if (i < np) {
            Acc ex = acc_add(carry, inc);
            ex.lights -= v.lights;
            ex.def -= v.def;
            ex.sur -= v.sur;
            part[i] = ex;
        }
        const Acc last = sh[kBlock - 1];
        __syncthreads();
        carry = acc_add(carry, last);
        (void)last;
    }
    if (threadIdx.x == 0) part[np] = carry;
}

// Scan phase 3: every item's exclusive prefix -> light / heavy lists.
__global__ __launch_bounds__(kBlock) void k_scan_scatter(const double* __restrict__ p, uint32_t n, const Acc* part,
                                                         uint32_t* light_idx, double* light_start, double* light_def,
                                                         uint32_t* heavy_idx, double* heavy_start) {
    __shared__ Acc sh[kBlock];
    const uint32_t base = blockIdx.x * kTile + threadIdx.x * kPerThread;
    Acc own{0u, 0.0, 0.0};
    for (int k = 0; k < kPerThread; ++k) own = acc_add(own, item_acc(p, base + k, n));
    const Acc inc = block_scan(own, sh);
    Acc run = acc_add(part[blockIdx.x], inc);
    run.lights -= own.lights;
    run.def -= own.def;
    run.sur -= own.sur;
    for (int k = 0; k < kPerThread; ++k) {
        const uint32_t i = base + k;
        if (i >= n) break;
        const Acc a = item_acc(p, i, n);
        if (a.lights) {
            light_idx[run.lights] = i;
            light_start[run.lights] = run.def;
            light_def[run.lights] = a.def;
        } else {
            const uint32_t m = i - run.lights;
            heavy_idx[m] = i;
            heavy_start[m] = run.sur;
        }
        run = acc_add(run, a);
    }
}

// Last index k in [0, m) with a[k] <= x (or, with strict, a[k] < x); -1 if none.
__device__ __forceinline__ int last_at_most(const double* a, uint32_t m, double x, bool strict) {
    int lo = 0, hi = static_cast<int>(m);  // answer in [lo - 1, hi - 1]
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const bool ok = strict ? a[mid] < x : a[mid] <= x;
        if (ok) lo = mid + 1;
        else hi = mid;
    }
    return lo - 1;
}

__global__ __launch_bounds__(kBlock) void k_assign_lights(const double* __restrict__ p, const Acc* totals,
                                                          const uint32_t* light_idx, const double* light_start,
                                                          const uint32_t* heavy_idx, const double* heavy_start,
                                                          uint32_t n, float2* table) {
    const uint32_t L = totals->lights, H = n - L;
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= L) return;
    const uint32_t i = light_idx[k];
    if (H == 0) {  // only possible through rounding: keep the item
        table[i] = make_float2(1.0f, __uint_as_float(i));
        return;
    }
    int m = last_at_most(heavy_start, H, light_start[k], false);
    if (m < 0) m = 0;
    table[i] = make_float2(static_cast<float>(p[i]), __uint_as_float(heavy_idx[m]));
}

__global__ __launch_bounds__(kBlock) void k_assign_heavies(const Acc* totals, const uint32_t* light_idx,
                                                           const double* light_start, const double* light_def,
                                                           const uint32_t* heavy_idx, const double* heavy_start,
                                                           uint32_t n, float2* table) {
    const uint32_t L = totals->lights, H = n - L;
    const uint32_t m = blockIdx.x * kBlock + threadIdx.x;
    if (m >= H) return;
    const uint32_t i = heavy_idx[m];
    float prob = 1.0f;
    uint32_t alias = i;
    if (m + 1 < H && L > 0) {
        const double x = heavy_start[m + 1];
        const int k = last_at_most(light_start, L, x, true);
        if (k >= 0) {
            const double over = light_start[k] + light_def[k] - x;
            if (over > 0.0) {
                prob = static_cast<float>(fmin(fmax(1.0 - over, 0.0), 1.0));
                alias = heavy_idx[m + 1];
            }
        }
    }
    table[i] = make_float2(prob, __uint_as_float(alias));
}

// light_sampler.cpp:67-81: area * max(luminance(emissive * strength), 0) for
// faces whose material emits (material.h:83-86), 0 for the others.
__global__ __launch_bounds__(kBlock) void k_emissive_weights(const float4* __restrict__ vpos,
                                                             const uint4* __restrict__ faces, uint32_t nf,
                                                             const pt_material* __restrict__ mats, uint32_t nmat,
                                                             float* wts) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= nf) return;
    const uint4 f = faces[i];
    float w = 0.0f;
    if (f.w < nmat) {  // a negative material id reads as a large unsigned one
        const pt_material& m = mats[f.w];
        if (m.emissive_strength > 0.0f && (m.emissive.x > 0.0f || m.emissive.y > 0.0f || m.emissive.z > 0.0f)) {
            const float4 a = vpos[f.x], b = vpos[f.y], c = vpos[f.z];
            const float e1x = b.x - a.x, e1y = b.y - a.y, e1z = b.z - a.z;
            const float e2x = c.x - a.x, e2y = c.y - a.y, e2z = c.z - a.z;
            const float cx = e1y * e2z - e1z * e2y, cy = e1z * e2x - e1x * e2z, cz = e1x * e2y - e1y * e2x;
            const float area = 0.5f * sqrtf(cx * cx + cy * cy + cz * cz);
            const float s = m.emissive_strength;
            const float le = 0.2126f * (m.emissive.x * s) + 0.7152f * (m.emissive.y * s) + 0.0722f * (m.emissive.z * s);
            if (area > 0.0f) w = fmaxf(area * fmaxf(le, 0.0f), 0.0f);
        }
    }
    wts[i] = w;
}

__device__ __forceinline__ Cnt face_cnt(const float* w, uint32_t i, uint32_t n) {
    if (i >= n) return Cnt{0u, 0.0};
    const float v = w[i];
    return v > 0.0f ? Cnt{1u, static_cast<double>(v)} : Cnt{0u, 0.0};
}

__global__ __launch_bounds__(kBlock) void k_cnt_tiles(const float* __restrict__ w, uint32_t n, Cnt* part) {
    __shared__ Cnt sh[kBlock];
    const uint32_t base = blockIdx.x * kTile + threadIdx.x * kPerThread;
    Cnt v{0u, 0.0};
    for (int k = 0; k < kPerThread; ++k) v = acc_add(v, face_cnt(w, base + k, n));
    v = block_scan(v, sh);
    if (threadIdx.x == kBlock - 1) part[blockIdx.x] = v;
}

// Exclusive offsets of the tiles (one block); the grand total in part[np].
__global__ __launch_bounds__(kBlock) void k_cnt_parts(Cnt* part, uint32_t np) {
    __shared__ Cnt sh[kBlock];
    Cnt carry{0u, 0.0};
    for (uint32_t b0 = 0; b0 < np; b0 += kBlock) {
        const uint32_t i = b0 + threadIdx.x;
        const Cnt v = i < np ? part[i] : Cnt{0u, 0.0};
        const Cnt inc = block_scan(v, sh);
        if (i < np) part[i] = Cnt{carry.n + inc.n - v.n, carry.w + inc.w - v.w};
        const Cnt last = sh[kBlock - 1];
        __syncthreads();
        carry = acc_add(carry, last);
    }
    if (threadIdx.x == 0) part[np] = carry;
}

// Emitting faces in face order: their weights, face indices and cdf values
// (running weight / total; the last entry is 1, light_sampler.cpp:93-105).
__global__ __launch_bounds__(kBlock) void k_cnt_compact(const float* __restrict__ w, uint32_t n, const Cnt* part,
                                                        float* wc, uint32_t* face, float* cdf) {
    __shared__ Cnt sh[kBlock];
    const uint32_t base = blockIdx.x * kTile + threadIdx.x * kPerThread;
    Cnt own{0u, 0.0};
    for (int k = 0; k < kPerThread; ++k) own = acc_add(own, face_cnt(w, base + k, n));
    const Cnt inc = block_scan(own, sh);
    const Cnt tot = part[gridDim.x];
    Cnt run = acc_add(part[blockIdx.x], inc);
    run.n -= own.n;
    run.w -= own.w;
    for (int k = 0; k < kPerThread; ++k) {
        const uint32_t i = base + k;
        if (i >= n) break;
        const Cnt a = face_cnt(w, i, n);
        if (a.n) {
            run = acc_add(run, a);
            wc[run.n - 1] = w[i];
            face[run.n - 1] = i;
            cdf[run.n - 1] = run.n == tot.n ? 1.0f : static_cast<float>(run.w / tot.w);
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_emissive_refs(const uint32_t* __restrict__ face, const float* __restrict__ pmf,
                                                          const float* __restrict__ cdf, uint32_t m, uint4* refs,
                                                          float* tri_pmf) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= m) return;
    const uint32_t f = face[k];
    refs[k] = make_uint4(f, __float_as_uint(pmf[k]), __float_as_uint(cdf[k]), 0u);
    tri_pmf[f] = pmf[k];
}

template <class T>
hipError_t dmalloc(T*& p, size_t n) {
    return hipMalloc(reinterpret_cast<void**>(&p), sizeof(T) * (n ? n : 1));
}

// The alias build's working buffers, carved from one block of device memory
// (every piece 256-byte aligned).  Every element a kernel reads is written
// earlier in the same build, so a block needs no clearing between builds.
struct AliasScratch {
    double* part_sum;     // per-tile weight sums
    double* total;        // the weight sum
    Acc* part;            // per-tile (lights, deficits, surpluses), then their total
    double* p;            // n * w_i / W
    uint32_t* light_idx;  // light items in index order
    double* light_start;  // ... where each one's deficit starts
    double* light_def;    // ... and its size
    uint32_t* heavy_idx;  // heavy items in index order
    double* heavy_start;  // ... where each one's surplus starts
};

// Byte offsets of the pieces above, in that order, and the block size.
struct AliasLayout {
    size_t off[9];
    size_t bytes;
};
constexpr AliasLayout alias_layout(uint32_t n) {
    const size_t tiles = (static_cast<size_t>(n) + kTile - 1) / kTile;
    const size_t sizes[9] = {sizeof(double) * tiles, sizeof(double), sizeof(Acc) * (tiles + 1),
                             sizeof(double) * n,     sizeof(uint32_t) * n, sizeof(double) * n,
                             sizeof(double) * n,     sizeof(uint32_t) * n, sizeof(double) * n};
    AliasLayout L{};
    size_t o = 0;
    for (int k = 0; k < 9; ++k) {
        o = (o + 255) & ~static_cast<size_t>(255);
        L.off[k] = o;
        o += sizes[k];
    }
    L.bytes = o;
    return L;
}

AliasScratch carve(uint32_t n, unsigned char* base) {
    const AliasLayout L = alias_layout(n);
    AliasScratch a;
    a.part_sum = reinterpret_cast<double*>(base + L.off[0]);
    a.total = reinterpret_cast<double*>(base + L.off[1]);
    a.part = reinterpret_cast<Acc*>(base + L.off[2]);
    a.p = reinterpret_cast<double*>(base + L.off[3]);
    a.light_idx = reinterpret_cast<uint32_t*>(base + L.off[4]);
    a.light_start = reinterpret_cast<double*>(base + L.off[5]);
    a.light_def = reinterpret_cast<double*>(base + L.off[6]);
    a.heavy_idx = reinterpret_cast<uint32_t*>(base + L.off[7]);
    a.heavy_start = reinterpret_cast<double*>(base + L.off[8]);
    return a;
}

// The build proper on a carved scratch block.
hipError_t alias_build(const float* weights, uint32_t n, float* pmf, float2* alias, const AliasScratch& a,
                       hipStream_t stream, std::string* err) {
    const uint32_t tiles = (n + kTile - 1) / kTile;
    const uint32_t blocks = (n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_sum_tiles, dim3(tiles), dim3(kBlock), 0, stream, weights, n, a.part_sum);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(kBlock), 0, stream, a.part_sum, tiles, a.total);
    hipError_t e = hipGetLastError();
    double W = 0.0;
    if (e == hipSuccess) e = hipMemcpyAsync(&W, a.total, sizeof(double), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return e;
    if (!(W > 0.0)) {
        if (err) *err = "alias table: the weights sum to zero";
        return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(k_scale, dim3(blocks), dim3(kBlock), 0, stream, weights, n, a.total, a.p, pmf);
    hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(kBlock), 0, stream, a.p, n, a.part);
    hipLaunchKernelGGL(k_scan_parts, dim3(1), dim3(kBlock), 0, stream, a.part, tiles);
    hipLaunchKernelGGL(k_scan_scatter, dim3(tiles), dim3(kBlock), 0, stream, a.p, n, a.part, a.light_idx,
                       a.light_start, a.light_def, a.heavy_idx, a.heavy_start);
    hipLaunchKernelGGL(k_assign_lights, dim3(blocks), dim3(kBlock), 0, stream, a.p, a.part + tiles, a.light_idx,
                       a.light_start, a.heavy_idx, a.heavy_start, n, alias);
    hipLaunchKernelGGL(k_assign_heavies, dim3(blocks), dim3(kBlock), 0, stream, a.part + tiles, a.light_idx,
                       a.light_start, a.light_def, a.heavy_idx, a.heavy_start, n, alias);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    return e;
}

// Scratch of the build that allocates nothing: one array per GPU, held until
// the process ends, shared by every call on that GPU.
__device__ __align__(256) unsigned char scratch[alias_layout(PT_ALIAS_NOALLOC_MAX_WEIGHTS).bytes];

// scratch's device address on each GPU, looked up on the first call there.
std::mutex g_scratch_mu;
std::vector<unsigned char*> g_scratch_ptr;

hipError_t scratch_on_current_device(unsigned char** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (static_cast<size_t>(dev) >= g_scratch_ptr.size()) g_scratch_ptr.resize(static_cast<size_t>(dev) + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = static_cast<unsigned char*>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

}  // namespace

hipError_t build_alias_device(const float* weights, uint32_t n, float* pmf, float2* alias, hipStream_t stream,
                              std::string* err) {
    if (n == 0) return hipSuccess;
    unsigned char* block = nullptr;
    hipError_t e = dmalloc(block, alias_layout(n).bytes);
    if (e != hipSuccess) return e;
    e = alias_build(weights, n, pmf, alias, carve(n, block), stream, err);
    (void)hipStreamSynchronize(stream);
    (void)hipFree(block);
    return e;
}

hipError_t build_alias_device_noalloc(const float* weights, uint32_t n, float* pmf, float2* alias, hipStream_t stream,
                                      std::string* err) {
    if (n == 0) return hipSuccess;
    if (n > PT_ALIAS_NOALLOC_MAX_WEIGHTS) {
        if (err)
            *err = "alias table without allocation: " + std::to_string(n) + " weights exceed the maximum of " +
                   std::to_string(PT_ALIAS_NOALLOC_MAX_WEIGHTS);
        return hipErrorInvalidValue;
    }
    unsigned char* block = nullptr;
    hipError_t e = scratch_on_current_device(&block);
    if (e != hipSuccess) return e;
    return alias_build(weights, n, pmf, alias, carve(n, block), stream, err);
}

hipError_t build_env_tables_device(const float4* env, int w, int h, float* pmf, float2* alias, hipStream_t stream,
                                   std::string* err) {
    if (w <= 0 || h <= 0 || !env) return hipSuccess;
    const uint32_t n = static_cast<uint32_t>(w) * static_cast<uint32_t>(h);
    float* wts = nullptr;
    hipError_t e = dmalloc(wts, n);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_env_weights, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, env, w, h, wts);
    e = hipGetLastError();
    if (e == hipSuccess) e = build_alias_device(wts, n, pmf, alias, stream, err);
    (void)hipStreamSynchronize(stream);
    (void)hipFree(wts);
    return e;
}

hipError_t build_emissive_tables_device(const float4* vpos, const uint4* faces, uint32_t nf, const pt_material* mats,
                                        uint32_t nmat, float* tri_pmf, uint32_t* count, uint4** emissive,
                                        float2** alias, hipStream_t stream, std::string* err) {
    *count = 0;
    *emissive = nullptr;
    *alias = nullptr;
    if (nf == 0) return hipSuccess;
    const uint32_t tiles = (nf + kTile - 1) / kTile;
    const uint32_t blocks = (nf + kBlock - 1) / kBlock;
    float *w = nullptr, *wc = nullptr, *cdf = nullptr, *pmf = nullptr;
    uint32_t* face = nullptr;
    Cnt* part = nullptr;
    uint4* refs = nullptr;
    float2* table = nullptr;
    hipError_t e = hipSuccess;
    auto ok = [&](hipError_t r) {
        if (e == hipSuccess) e = r;
        return e == hipSuccess;
    };
    Cnt tot{0u, 0.0};
    if (ok(dmalloc(w, nf)) && ok(dmalloc(part, tiles + 1)) && ok(hipMemsetAsync(tri_pmf, 0, sizeof(float) * nf, stream))) {
        hipLaunchKernelGGL(k_emissive_weights, dim3(blocks), dim3(kBlock), 0, stream, vpos, faces, nf, mats, nmat, w);
        hipLaunchKernelGGL(k_cnt_tiles, dim3(tiles), dim3(kBlock), 0, stream, w, nf, part);
        hipLaunchKernelGGL(k_cnt_parts, dim3(1), dim3(kBlock), 0, stream, part, tiles);
        if (ok(hipGetLastError()) && ok(hipMemcpyAsync(&tot, part + tiles, sizeof(Cnt), hipMemcpyDeviceToHost, stream)))
            ok(hipStreamSynchronize(stream));
    }
    const uint32_t m = tot.n;
    if (e == hipSuccess && m > 0 && ok(dmalloc(wc, m)) && ok(dmalloc(cdf, m)) && ok(dmalloc(pmf, m)) &&
        ok(dmalloc(face, m)) && ok(dmalloc(refs, m)) && ok(dmalloc(table, m))) {
        hipLaunchKernelGGL(k_cnt_compact, dim3(tiles), dim3(kBlock), 0, stream, w, nf, part, wc, face, cdf);
        if (ok(hipGetLastError()) && ok(build_alias_device(wc, m, pmf, table, stream, err))) {
            hipLaunchKernelGGL(k_emissive_refs, dim3((m + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, face, pmf,
                               cdf, m, refs, tri_pmf);
            if (ok(hipGetLastError())) ok(hipStreamSynchronize(stream));
        }
    }
    (void)hipStreamSynchronize(stream);
    void* bufs[] = {w, wc, cdf, pmf, face, part};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (e == hipSuccess && m > 0) {
        *count = m;
        *emissive = refs;
        *alias = table;
    } else {
        if (refs) (void)hipFree(refs);
        if (table) (void)hipFree(table);
    }
    return e;
}

}  // namespace ptd

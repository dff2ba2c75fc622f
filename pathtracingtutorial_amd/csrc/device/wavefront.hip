// Wavefront form of the path tracer for gfx950.
//
// The estimator is the megakernel's (dev_integrator.h: same bounce() code,
// same random numbers per path); only the schedule changes:
//
//  * A batch renders one or many jobs (scene + camera + resolution + spp +
//    film).  Job j owns paths [O_j, O_j + W_j*H_j*spf_j) of a global path
//    index, path O_j + pixel*spf_j + sample; the index space is processed in
//    chunks of C paths (C = chunk_spp x all jobs' pixels), so scenes of any
//    size and resolution share the same persistent kernels and queues.
//  * Path state lives in HBM as SoA float4 arrays (ray origin, direction +
//    last pdf, throughput + depth/flags, radiance, sampler state + job, hit).
//  * Per bounce: a persistent closest-hit kernel over the ray queue (lean:
//    traversal only, LDS stack), an optional counting sort of the queue by
//    material, a persistent shading kernel that appends next rays and NEE
//    shadow rays to queues, and persistent any-hit kernels over the area,
//    environment and punctual shadow queues.  Area and environment shadow
//    rays go to separate queues processed by separate launches, so each path
//    gains at most one contribution per launch and no atomics are needed.
//  * Work is grabbed per wave (64 items, one atomic) from 8 queue segments,
//    one per XCD, so each XCD's L2 sees contiguous (coherent) rays; a wave
//    steals from the other segments once its own is empty.
//  * A pixel's samples are summed in sample order across chunks, so the image
//    is bitwise the megakernel's sum (up to FMA contraction differences).
//  * The bounce loop of a chunk is captured once in a hipGraph (job table and
//    chunk index live in device memory) and replayed per chunk.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "adjoint_common.h"
#include "dev_integrator.h"
#include "wavefront.h"

namespace ptd {

namespace {

constexpr int kSeg = 8;
// Blocks per segment of the counting sorts (their histograms are per block):
// 128 for the material sort (1080p frame: sort 4.8 -> 2.4 ms with 32), 64 for
// the replay's splat records (whose scan covers up to 4096 tiles per part).
constexpr int kSortParts = 128;
constexpr int kSplatSortParts = 64;
constexpr uint32_t kMaxSortBins = 8192;
// The replay hands the environment gradients of its escapes after the first
// bounce and of its unoccluded environment shadow rays to a splat stage
// instead of adding them with device atomics one texel tap at a time (their
// directions are random, so a wave's taps rarely share a texel): each one's
// taps are recorded (a path makes at most one such record per bounce), the
// bounce's records are counting-sorted by the kSplatTileW x kSplatTileH tile
// of their first tap, and one block per (segment, tile) sums the tile's
// records in LDS and adds the sums to the gradient with one atomic per
// touched texel channel.
constexpr int kSplatTileW = 32, kSplatTileH = 16;
constexpr uint32_t kMaxSplatTiles = 4096;
constexpr uint32_t kSplatParts = 4;  // blocks per (segment, tile), each a quarter of its records
// control words
constexpr int C_CHUNK = 33;  // chunk index within the batch
constexpr int C_CUR_EXT = 40, C_CUR_SHADE = 48, C_CUR_SA = 56, C_CUR_SE = 64, C_CUR_SP = 72, C_CUR_MISS = 80;
constexpr int C_MISS0 = 88;  // [8] first item of each segment's escape bin in the material-sorted queue
// per cursor set (6 of them, in the order above): a bit mask of drained
// segments, each mask on a 128-byte line of its own
constexpr int C_DRAIN = 128, C_DRAIN_STRIDE = 32;
// The cursors themselves (C_CUR_* name a set): each on a 128-byte line of
// its own, cursor k of set C_CUR_X at C_CURSORS + (C_CUR_X - C_CUR_EXT + k)
// lines.  The waves of the whole device add to them; on one shared line the
// eight segments' atomics would queue behind one another.
constexpr int C_LINE = 32;
constexpr int C_CURSORS = C_DRAIN + 6 * C_DRAIN_STRIDE;
// The queues' segment counts, which every wave of the shading kernels adds
// to, each on a line of its own too: count k of a set at base + k * C_LINE.
constexpr int C_COUNTS = C_CURSORS + 6 * 8 * C_LINE;
constexpr int C_QCNT = C_COUNTS;                      // [2][8] ray queue segment counts
constexpr int C_SACNT = C_COUNTS + 2 * 8 * C_LINE;  // [8] area shadow queue segment counts
constexpr int C_SECNT = C_COUNTS + 3 * 8 * C_LINE;  // [8] environment shadow queue segment counts
constexpr int C_SP = C_COUNTS + 4 * 8 * C_LINE;     // punctual shadow queue count (one flat queue)
constexpr int C_ESCNT = C_SP + C_LINE;              // [8] replay: escape records per segment
constexpr int C_WORDS = C_ESCNT + 8 * C_LINE;

// A job as the kernels see it (device memory, one entry per job).
struct JobDesc {
    DevScene scene;
    RenderParams params;
    float4* film_new;
    float4* film_acc;
    const float4* img_grad;          // adjoint: d loss / d batch mean image
    float4* replay;                  // adjoint: batch mean of the replay pass, or null
    unsigned long long path_offset;  // first global path
    uint32_t pixel_offset;           // first global pixel (per-pixel sums)
    uint32_t npix;
};

// A path's ray, throughput and sampler state: one 64-byte record, so the
// shading kernel, which visits paths in material order (scattered), fetches
// one line per path instead of one per field.
struct PathRec {
    float4 o;   // xyz origin
    float4 d;   // xyz direction, w = last bsdf pdf
    float4 tp;  // xyz throughput, w = bits: depth | last_delta << 31
    uint4 rng;  // x job, y Sobol index, z Sobol dim or PCG state, w pixel hash
};

struct WfState {
    PathRec* path;
    float4* lo;   // xyz radiance (apart: the shadow kernels and the per-pixel sums touch only it)
    float4* hit;  // t, u, v, triangle bits
    uint32_t* q[2];
    uint32_t* sorted;
    uint32_t* qkey;       // material key of each ray-queue item, in queue order (written by the closest-hit kernel)
    uint32_t key_bins;    // material bins + the miss bin when sorting, else 0
    float4 *sa_o, *sa_d, *sa_c;  // area shadow queue: origin + tmax, direction + path bits, contribution
    float4 *se_o, *se_d, *se_c;  // environment shadow queue
    float4 *sp_o, *sp_d, *sp_c;  // punctual shadow queue
    uint32_t* blockhist;         // material sort: [segment][part][bin] counts, then offsets
    float4* sum;                 // per global pixel sum over the batch
    uint32_t* ctrl;
    uint32_t* spill;  // traversal stack entries beyond the LDS part, [level][thread of the launch]
    // adjoint (replay) state
    float4* adj_L;                // primal radiance of each path
    float4* adj_g;                // d loss / d path radiance (zero above the clamp)
    float4* pend;                 // kPend per path: the last continuation's sensitivities
    float4 *sa_x, *se_x, *sp_x;   // kPayload per shadow entry: its NEE sensitivities
    float4* esc_rec;              // replay: environment-gradient records (u, v, coefficient x, y)
    float* esc_cz;                // their coefficient's z
    uint32_t* esc_key;            // their splat tile
    float4* esc_srec;             // the records of each segment in tile order
    float* esc_scz;
    uint32_t* esc_hist;           // [segment][part][tile] counts, then offsets, of that order
    uint32_t env_tiles_x;         // splat tiles across the environment
    uint32_t env_tiles;           // splat tiles in all (0: escapes add their gradients directly)
    // each path's radiance sum by global path index, kept from a forward batch
    // for the adjoint of the same batch: kept_mode 1 = the forward writes it
    // (per-chunk finalize), 2 = the adjoint's seed reads it and the chunk's
    // primal pass is not run, 0 = neither
    float4* kept;
    uint32_t kept_mode;
    const AdjointGrads* grads;    // parameter gradient buffers (device copy)
    const JobDesc* jobs;
    DevCounters* shade_ctr;  // the shading kernel's event counters (counted runs only, else null)
    uint32_t njobs;
    uint32_t capacity, sp_capacity;
    uint32_t chunk_paths;              // C
    uint32_t grab_extend, grab_shade, grab_shadow;  // queue items a wave takes per atomic
    unsigned long long total_paths;    // all jobs' paths of the batch
    uint32_t total_pixels;
};

__device__ __forceinline__ int lane_id() { return static_cast<int>(threadIdx.x & 63); }

__device__ __forceinline__ uint32_t seg_base(uint32_t n, int k) {
    return static_cast<uint32_t>((static_cast<unsigned long long>(n) * static_cast<unsigned>(k)) / kSeg);
}

// Paths of the current chunk and its first global path.
__device__ __forceinline__ uint32_t chunk_n(const WfState& w, unsigned long long& c0) {
    c0 = static_cast<unsigned long long>(w.ctrl[C_CHUNK]) * w.chunk_paths;
    const unsigned long long left = w.total_paths > c0 ? w.total_paths - c0 : 0ull;
    return static_cast<uint32_t>(left < w.chunk_paths ? left : w.chunk_paths);
}

// Job of a global path / global pixel (the job table is sorted by both).
__device__ __forceinline__ uint32_t job_of_path(const WfState& w, unsigned long long g) {
    uint32_t lo = 0, hi = w.njobs;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (w.jobs[mid].path_offset <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}
__device__ __forceinline__ uint32_t job_of_pixel(const WfState& w, uint32_t gp) {
    uint32_t lo = 0, hi = w.njobs;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (w.jobs[mid].pixel_offset <= gp) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint32_t path_job(const WfState& w, uint32_t p) {
    return w.path[p].rng.x;
}

// Wave-aggregated append: every lane of the wave must call it.
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool pred) {
    unsigned long long mask = __ballot(pred);
    uint32_t base = 0;
    if (mask) {
        int leader = __ffsll(static_cast<long long>(mask)) - 1;
        if (lane_id() == leader) base = atomicAdd(counter, static_cast<uint32_t>(__popcll(mask)));
        base = __shfl(base, leader, 64);
    }
    unsigned long long below = (lane_id() == 0) ? 0ull : (mask & ((1ull << lane_id()) - 1ull));
    return base + static_cast<uint32_t>(__popcll(below));
}

// A wave's walk over a queue split in kSeg segments (one per XCD): it grabs
// `step` items (a multiple of 64) with one atomic on the segment's cursor,
// its own segment first, then the others, and hands them out 64 at a time.
// Segment k holds [base(k), base(k) + counts[k]) (counts == null: a flat
// queue of n items in kSeg equal ranges).  Grabbing several rows of 64 per
// atomic keeps the same-address atomics on the 8 cursors, which the device
// serializes, off the critical path.  Draining is cheap too: an empty
// segment is skipped without an atomic, and the one wave whose grab is the
// first past a segment's end sets the segment's bit in the drained mask,
// which waves read before stealing from another segment; otherwise every
// wave of the persistent grid would end with one atomic per segment.
struct QueueWalk {
    uint32_t* cursors;
    uint32_t* drained;
    const uint32_t* counts;
    const uint32_t* starts;  // per segment: its first item instead of base(k) (or null)
    const uint32_t* stops;   // per segment: its end instead of base(k) + counts[k] (or null)
    uint32_t n, step;
    int seg, tried;
    uint32_t cur, lim;  // the wave's current grab [cur, lim)

    __device__ __forceinline__ QueueWalk(uint32_t* ctrl, int cur_slot, const uint32_t* cnt, uint32_t items,
                                         uint32_t grab, const uint32_t* seg_starts = nullptr,
                                         const uint32_t* seg_stops = nullptr)
        : cursors(ctrl + C_CURSORS + (cur_slot - C_CUR_EXT) * C_LINE),
          drained(ctrl + C_DRAIN + C_DRAIN_STRIDE * ((cur_slot - C_CUR_EXT) / kSeg)),
          counts(cnt), starts(seg_starts), stops(seg_stops), n(items), step(grab),
          seg(static_cast<int>(blockIdx.x % kSeg)), tried(0), cur(0), lim(0) {}

    __device__ __forceinline__ void skip() {
        seg = (seg + 1) % kSeg;
        ++tried;
    }

    // The next 64 items of the wave: lane item i, valid while i < end.
    // Returns false (for the whole wave) once every segment is drained.
    __device__ __forceinline__ bool next(uint32_t& i, uint32_t& end) {
        while (cur >= lim) {
            if (tried > 0 && tried < kSeg) {  // stealing: leave out segments known to be drained
                uint32_t m = 0;
                if (lane_id() == 0) m = __hip_atomic_load(drained, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                m = __shfl(m, 0, 64);
                while (tried < kSeg && ((m >> seg) & 1u)) skip();
            }
            if (tried >= kSeg) return false;
            const uint32_t b = seg_base(n, seg);
            const uint32_t s0 = starts ? starts[seg] : b;
            const uint32_t s1 = stops ? stops[seg] : (counts ? b + counts[seg * C_LINE] : seg_base(n, seg + 1));
            if (s1 <= s0) {
                skip();
                continue;
            }
            uint32_t off = 0;
            if (lane_id() == 0) {
                off = atomicAdd(&cursors[seg * C_LINE], step);
                if (off >= s1 - s0 && off < s1 - s0 + step) atomicOr(drained, 1u << seg);
            }
            off = __shfl(off, 0, 64);
            if (s0 + off < s1) {
                cur = s0 + off;
                lim = min(s1, cur + step);
            } else {
                skip();
            }
        }
        i = cur + static_cast<uint32_t>(lane_id());
        end = lim;
        cur += 64;
        return true;
    }
};

__device__ __forceinline__ void load_path(const WfState& w, uint32_t p, PathState& st, uint4& r) {
    const PathRec& R = w.path[p];
    float4 o = R.o, d = R.d, tp = R.tp, lo = w.lo[p];
    st.o = f3(o);
    st.d = f3(d);
    st.last_pdf = d.w;
    st.tp = f3(tp);
    uint32_t bits = __float_as_uint(tp.w);
    st.depth = static_cast<int>(bits & 0xFFFFu);
    st.last_delta = (bits >> 31) != 0;
    st.lo = f3(lo);
    r = R.rng;
}

__device__ __forceinline__ void store_path(const WfState& w, uint32_t p, const PathState& st, uint4 r) {
    PathRec& R = w.path[p];
    R.o = make_float4(st.o.x, st.o.y, st.o.z, 0.0f);
    R.d = make_float4(st.d.x, st.d.y, st.d.z, st.last_pdf);
    uint32_t bits = static_cast<uint32_t>(st.depth) | (st.last_delta ? 0x80000000u : 0u);
    R.tp = make_float4(st.tp.x, st.tp.y, st.tp.z, __uint_as_float(bits));
    w.lo[p] = make_float4(st.lo.x, st.lo.y, st.lo.z, 0.0f);
    R.rng = r;
}

// store_path for a bounce that started from st.lo = 0: the radiance it added
// (an emitter hit or an escape, no more than once per bounce) goes onto the
// path's sum, and a bounce that added none leaves the sum untouched.  Each
// path is in a queue once, and the shadow kernels that also add to the sum
// run in other launches.  The record of a path that ends here is dead (the
// per-pixel sums and the adjoint's seed read only the sum; a replay starts
// from fresh camera rays): it is written only for a continuing path.
__device__ __forceinline__ void store_path_added(const WfState& w, uint32_t p, const PathState& st, uint4 r,
                                                 bool cont) {
    if (cont) {
        PathRec& R = w.path[p];
        R.o = make_float4(st.o.x, st.o.y, st.o.z, 0.0f);
        R.d = make_float4(st.d.x, st.d.y, st.d.z, st.last_pdf);
        uint32_t bits = static_cast<uint32_t>(st.depth) | (st.last_delta ? 0x80000000u : 0u);
        R.tp = make_float4(st.tp.x, st.tp.y, st.tp.z, __uint_as_float(bits));
        R.rng = r;
    }
    if (!is_zero(st.lo)) {
        const float4 l = w.lo[p];
        w.lo[p] = make_float4(l.x + st.lo.x, l.y + st.lo.y, l.z + st.lo.z, 0.0f);
    }
}

__device__ __forceinline__ Sampler sampler_of(int type, uint4 r) {
    Sampler s;
    s.type = static_cast<uint32_t>(type);
    s.index = r.y;
    s.dim = r.z;
    s.state = r.z;
    s.pix_hash = r.w;
    return s;
}
__device__ __forceinline__ uint4 sampler_state(const Sampler& s, uint4 r) {
    r.z = s.type == 1 ? s.dim : s.state;
    return r;
}

template <bool COUNT, bool SHADOW>
__device__ __forceinline__ void flush(DevCounters* ctr, unsigned long long rays, unsigned long long nodes,
                                      unsigned long long tris, unsigned long long overflow) {
    if (!COUNT) return;
    rays = wave_sum_u64(rays);
    nodes = wave_sum_u64(nodes);
    tris = wave_sum_u64(tris);
    overflow = wave_sum_u64(overflow);
    if (lane_id() == 0) {
        atomicAdd(SHADOW ? &ctr->shadow_rays : &ctr->closest_rays, rays);
        atomicAdd(&ctr->node_visits, nodes);
        atomicAdd(&ctr->tri_tests, tris);
        if (SHADOW) {
            atomicAdd(&ctr->shadow_node_visits, nodes);
            atomicAdd(&ctr->shadow_tri_tests, tris);
        }
        if (overflow) atomicAdd(&ctr->stack_overflows, overflow);
    }
}

// ---------------------------------------------------------------- kernels

__global__ void k_wf_batch_begin(uint32_t* ctrl) {
    if (threadIdx.x == 0) ctrl[C_CHUNK] = 0;
}

// Ray queue 0 of a chunk: every path, segment k = [base(k), base(k+1)).
__global__ void k_wf_chunk_begin(WfState w) {
    const int k = static_cast<int>(threadIdx.x);
    unsigned long long c0;
    const uint32_t n = chunk_n(w, c0);
    if (k < kSeg) w.ctrl[C_QCNT + k * C_LINE] = seg_base(n, k + 1) - seg_base(n, k);
}

__global__ void k_wf_chunk_end(uint32_t* ctrl) {
    if (threadIdx.x == 0) ctrl[C_CHUNK] += 1;
}

// Camera rays of the chunk: global path g = O_j + pixel * spf_j + sample.
__global__ void k_wf_generate(WfState w) {
    unsigned long long c0;
    const uint32_t n = chunk_n(w, c0);
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const unsigned long long g = c0 + p;
    const uint32_t j = job_of_path(w, g);
    const JobDesc& job = w.jobs[j];
    const RenderParams& rp = job.params;
    const unsigned long long local = g - job.path_offset;
    const uint32_t spf = static_cast<uint32_t>(rp.spf);
    const uint32_t pixel = static_cast<uint32_t>(local / spf), sample = static_cast<uint32_t>(local % spf);
    int x = static_cast<int>(pixel % static_cast<uint32_t>(rp.width));
    int y = static_cast<int>(pixel / static_cast<uint32_t>(rp.width));
    Sampler rng{};
    if (rp.sampler == 1) rng.init_sobol(x, y, rp.frame_count * spf + sample, rp.seed);
    else rng.init_pcg(x, y, rp.frame_count, sample, rp.seed);
    PathState st;
    begin_path(rp, x, y, rng, st);
    uint4 r = make_uint4(j, rng.index, 0u, rng.pix_hash);
    r = sampler_state(rng, r);
    store_path(w, p, st, r);
    w.q[0][p] = p;
}

// Clears the output queue's and the shadow queues' counts and all cursors.
__global__ void k_wf_bounce_begin(uint32_t* ctrl, int out_slot) {
    uint32_t t = threadIdx.x;
    if (t < kSeg) {
        ctrl[C_QCNT + (out_slot * kSeg + t) * C_LINE] = 0;
        ctrl[C_SACNT + t * C_LINE] = 0;
        ctrl[C_SECNT + t * C_LINE] = 0;
        ctrl[C_ESCNT + t * C_LINE] = 0;
    }
    if (t == 0) ctrl[C_SP] = 0;
    // (the cursors, the drained masks and the escape-bin starts; not the
    // counts, which hold the input queue)
    for (uint32_t k = C_CUR_EXT + t; k < C_COUNTS; k += blockDim.x) ctrl[k] = 0;
}

// Traversal stack entries of the binary or the 4-wide BVH.
template <bool WIDE>
constexpr int stack_total() {
    return WIDE ? kWideStack : kStackSize;
}

template <int LDS, int BLOCK, int TOTAL = kStackSize>
__device__ __forceinline__ TraceStack<LDS, TOTAL> thread_stack(const WfState& w, uint32_t* lds) {
    return TraceStack<LDS, TOTAL>{lds + threadIdx.x, BLOCK, w.spill + blockIdx.x * BLOCK + threadIdx.x,
                                  gridDim.x * BLOCK};
}

template <bool COUNT, int BLOCK, int LDS, bool WIDE>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_wf_extend(WfState w, int in_slot,
                                                                                                DevCounters* ctr) {
    __shared__ uint32_t stack[LDS * BLOCK];
    const auto sk = thread_stack<LDS, BLOCK, stack_total<WIDE>()>(w, stack);
    unsigned long long c0;
    const uint32_t n = chunk_n(w, c0);
    const uint32_t* counts = w.ctrl + C_QCNT + in_slot * kSeg * C_LINE;
    const uint32_t* q = w.q[in_slot];
    QueueWalk qw(w.ctrl, C_CUR_EXT, counts, n, w.grab_extend);
    uint32_t i, end;
    TraceCounters cnt{0, 0};
    uint32_t ovf = 0;
    unsigned long long rays = 0;
    // One job: its scene is the same for every lane, so the traversal's base
    // addresses stay in scalar registers (a uniform branch picks that copy).
    const bool one_job = w.njobs == 1;
    while (qw.next(i, end)) {
        if (i < end) {
            const uint32_t p = q[i];
            const float4 o = w.path[p].o, d = w.path[p].d;
            Hit h;
            uint32_t mat = 0;
            if (one_job) {
                const DevScene& s = w.jobs[0].scene;
                h = trace_closest_v<WIDE, COUNT>(s, f3(o), f3(d), sk, cnt, ovf);
                if (w.key_bins && h.tri != 0xFFFFFFFFu) mat = gld_u32(&s.faces[h.tri].w);
            } else {
                const DevScene& s = w.jobs[path_job(w, p)].scene;
                h = trace_closest_v<WIDE, COUNT>(s, f3(o), f3(d), sk, cnt, ovf);
                if (w.key_bins && h.tri != 0xFFFFFFFFu) mat = gld_u32(&s.faces[h.tri].w);
            }
            w.hit[p] = make_float4(h.t, h.u, h.v, __uint_as_float(h.tri));
            // the sort's key, stored in queue order so the sort reads it contiguously
            if (w.key_bins) w.qkey[i] = h.tri == 0xFFFFFFFFu ? w.key_bins - 1 : min(mat, w.key_bins - 2);
            rays++;
        }
    }
    flush<COUNT, false>(ctr, rays, cnt.nodes, cnt.tris, ovf);
}

// ---- material sort: a counting sort within each segment -------------------
// Block j sorts part j / kSeg of segment j % kSeg: per-block histograms in
// LDS, one scan per segment over (bin, part), then a scatter that ranks the
// block's items with LDS atomics.  Misses go to the last bin; the key is the
// material index (of whichever job the path belongs to), written per queue
// item by the closest-hit kernel (qkey).

__device__ __forceinline__ void sort_part_range(uint32_t n, const uint32_t* counts, int seg, int part,
                                                uint32_t parts, uint32_t& b0, uint32_t& b1) {
    const uint32_t base = seg_base(n, seg), cnt = counts[seg * C_LINE];
    b0 = base + static_cast<uint32_t>((static_cast<unsigned long long>(cnt) * part) / parts);
    b1 = base + static_cast<uint32_t>((static_cast<unsigned long long>(cnt) * (part + 1)) / parts);
}

// What a counting sort within the segments works on: keys in queue order,
// the per-(segment, part) histograms, the output, the items it moves (null:
// the entries' own indices), the segment counts, for the material sort where
// each segment's escape bin starts, and the blocks (parts) per segment.
struct SortSpec {
    const uint32_t* key;
    uint32_t* hist;
    uint32_t* out;
    const uint32_t* src;
    const uint32_t* counts;
    uint32_t* miss0;
    uint32_t parts;
};

__global__ __launch_bounds__(256) void k_wf_sort_count(WfState w, SortSpec a, uint32_t nbins) {
    extern __shared__ uint32_t hist[];
    const int seg = static_cast<int>(blockIdx.x % kSeg), part = static_cast<int>(blockIdx.x / kSeg);
    for (uint32_t b = threadIdx.x; b < nbins; b += blockDim.x) hist[b] = 0;
    __syncthreads();
    unsigned long long c0;
    const uint32_t n = chunk_n(w, c0);
    uint32_t b0, b1;
    sort_part_range(n, a.counts, seg, part, a.parts, b0, b1);
    for (uint32_t i = b0 + threadIdx.x; i < b1; i += blockDim.x) atomicAdd(&hist[a.key[i]], 1u);
    __syncthreads();
    uint32_t* out = a.hist + (static_cast<size_t>(seg) * a.parts + part) * nbins;
    for (uint32_t b = threadIdx.x; b < nbins; b += blockDim.x) out[b] = hist[b];
}

// One block of T threads per segment: offsets in (bin, part) order, starting
// at the segment's base, written over the counts.  Each thread sums a run of
// entries; the runs' sums are scanned in LDS (log2 T steps).
template <int T>
__global__ __launch_bounds__(T) void k_wf_sort_scan(WfState w, SortSpec a, uint32_t nbins) {
    __shared__ uint32_t part_sum[T];
    const int seg = static_cast<int>(blockIdx.x);
    unsigned long long c0;
    const uint32_t n = chunk_n(w, c0);
    const uint32_t P = a.parts;
    uint32_t* h = a.hist + static_cast<size_t>(seg) * P * nbins;
    const uint32_t E = nbins * P;  // entry e = bin * parts + part
    const uint32_t t = threadIdx.x;
    const uint32_t e0 = static_cast<uint32_t>((static_cast<unsigned long long>(E) * t) / T);
    const uint32_t e1 = static_cast<uint32_t>((static_cast<unsigned long long>(E) * (t + 1)) / T);
    auto at = [&](uint32_t e) -> uint32_t& { return h[(e % P) * nbins + e / P]; };
    uint32_t sum = 0;
    for (uint32_t e = e0; e < e1; ++e) sum += at(e);
    part_sum[t] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < static_cast<uint32_t>(T); off <<= 1) {  // inclusive scan
        const uint32_t v = t >= off ? part_sum[t - off] : 0u;
        __syncthreads();
        part_sum[t] += v;
        __syncthreads();
    }
    uint32_t acc = seg_base(n, seg) + (t > 0 ? part_sum[t - 1] : 0u);
    const uint32_t e_miss = (nbins - 1) * P;  // the escape bin (the last), part 0
    for (uint32_t e = e0; e < e1; ++e) {
        uint32_t v = at(e);
        if (e == e_miss && a.miss0) a.miss0[seg] = acc;
        at(e) = acc;
        acc += v;
    }
}

__global__ __launch_bounds__(256) void k_wf_sort_scatter(WfState w, SortSpec a, uint32_t nbins) {
    extern __shared__ uint32_t off[];
    const int seg = static_cast<int>(blockIdx.x % kSeg), part = static_cast<int>(blockIdx.x / kSeg);
    const uint32_t* src = a.hist + (static_cast<size_t>(seg) * a.parts + part) * nbins;
    for (uint32_t b = threadIdx.x; b < nbins; b += blockDim.x) off[b] = src[b];
    __syncthreads();
    unsigned long long c0;
    const uint32_t n = chunk_n(w, c0);
    uint32_t b0, b1;
    sort_part_range(n, a.counts, seg, part, a.parts, b0, b1);
    for (uint32_t i = b0 + threadIdx.x; i < b1; i += blockDim.x)
        a.out[atomicAdd(&off[a.key[i]], 1u)] = a.src ? a.src[i] : i;
}

// Visibility policy of the wavefront: shadow rays go to the queues at once.
// Area and environment rays are appended with one atomic per wave (the lanes
// that reach this point; each path adds at most one entry per queue) to the
// segment the path's wave is working on; punctual rays with one atomic each.
struct DeferNee {
    static constexpr bool kDrawAhead = true;
    uint32_t light_samples = 0, entries = 0;  // events, for the counted runs
    uint32_t path;
    uint32_t base;  // first slot of the segment in the shadow queues
    uint32_t *sa_count, *se_count;
    float4 *sa_o, *sa_d, *sa_c, *se_o, *se_d, *se_c;
    uint32_t* sp_count;
    uint32_t sp_capacity;
    float4 *sp_o, *sp_d, *sp_c;
    __device__ __forceinline__ void direct(const DevScene&, PathState&, bool a, const ShadowRay& ra, F3 xa, bool e,
                                           const ShadowRay& re, F3 xe, const LightSample&, const LightSample&,
                                           const ShadeCtx&) {
        light_samples += 1;
        entries += (a ? 1u : 0u) + (e ? 1u : 0u);
        const uint32_t ka = base + wave_append(sa_count, a);
        if (a) {
            sa_o[ka] = make_float4(ra.o.x, ra.o.y, ra.o.z, ra.tmax);
            sa_d[ka] = make_float4(ra.d.x, ra.d.y, ra.d.z, __uint_as_float(path));
            sa_c[ka] = make_float4(xa.x, xa.y, xa.z, 0.0f);
        }
        const uint32_t ke = base + wave_append(se_count, e);
        if (e) {
            se_o[ke] = make_float4(re.o.x, re.o.y, re.o.z, re.tmax);
            se_d[ke] = make_float4(re.d.x, re.d.y, re.d.z, __uint_as_float(path));
            se_c[ke] = make_float4(xe.x, xe.y, xe.z, 0.0f);
        }
    }
    __device__ __forceinline__ void on_env_miss(const DevScene&, const PathState&, float) {}
    __device__ __forceinline__ void on_bsdf_sample(const DevScene&, const PathState&, const BsdfSample&,
                                                   const ShadeCtx&) {}
    __device__ __forceinline__ void punctual_lights(const DevScene& s, PathState& st, const ShadeCtx& x) {
        const F3 p = x.p, wo = x.wo, n = x.n;
        const MatEval<float>& m = *x.me;
        for (uint32_t i = 0; i < s.num_lights; ++i) {
            F3 wi, inten;
            float dist;
            if (!punctual_light_sample(s, i, p, n, wi, inten, dist)) continue;
            BsdfEvalT<float> b = eval_bsdf(m, wo, wi, n, true);
            F3 c = st.tp * (inten * f3(b.f.x, b.f.y, b.f.z));
            if (is_zero(c)) continue;
            uint32_t k = atomicAdd(sp_count, 1u);
            if (k >= sp_capacity) continue;  // cannot happen: capacity = paths x lights
            ShadowRay r = make_shadow(p, n, wi, dist);
            sp_o[k] = make_float4(r.o.x, r.o.y, r.o.z, r.tmax);
            sp_d[k] = make_float4(r.d.x, r.d.y, r.d.z, __uint_as_float(path));
            sp_c[k] = make_float4(c.x, c.y, c.z, 0.0f);
            entries += 1;
        }
    }
};

// Per-lane event counts of the shading kernel (pt_counters' shade_* fields).
struct ShadeEvents {
    uint32_t items = 0, env_lookups = 0, hits = 0, light_samples = 0, entries = 0, conts = 0, radiance_adds = 0;
    __device__ __forceinline__ void flush(DevCounters* c) const {
        const unsigned long long v[7] = {wave_sum_u64(items), wave_sum_u64(env_lookups), wave_sum_u64(hits),
                                         wave_sum_u64(light_samples), wave_sum_u64(entries), wave_sum_u64(conts),
                                         wave_sum_u64(radiance_adds)};
        if (lane_id() == 0) {
            atomicAdd(&c->shade_items, v[0]);
            atomicAdd(&c->shade_env_lookups, v[1]);
            atomicAdd(&c->shade_hits, v[2]);
            atomicAdd(&c->shade_light_samples, v[3]);
            atomicAdd(&c->shade_shadow_entries, v[4]);
            atomicAdd(&c->shade_continuations, v[5]);
            atomicAdd(&c->shade_radiance_adds, v[6]);
        }
    }
};

// One queue item of the shading kernel with its path record and hit as
// loaded from memory (item i of the wave's grab, valid while i < end).
struct ShadeItem {
    bool has = false;
    int seg = 0;
    uint32_t i = 0, end = 0, p = 0;
    float4 o, d, tp, hv;
    uint4 rng;
    __device__ __forceinline__ void fetch(const WfState& w) {
        const PathRec& R = w.path[p];
        o = R.o;
        d = R.d;
        tp = R.tp;
        rng = R.rng;
        hv = w.hit[p];
    }
    __device__ __forceinline__ void unpack(PathState& st) const {
        st.o = f3(o);
        st.d = f3(d);
        st.last_pdf = d.w;
        st.tp = f3(tp);
        const uint32_t bits = __float_as_uint(tp.w);
        st.depth = static_cast<int>(bits & 0xFFFFu);
        st.last_delta = (bits >> 31) != 0;
        st.lo = f3(0.0f);  // what this bounce adds; the path's sum stays in memory
    }
};

// WAVES > 0 asks the compiler for at least that many waves per SIMD (fewer
// registers, possibly spills): the shading kernel's occupancy knob.
// split: the queue is material-sorted and its escape bin is left to
// k_wf_shade_escape (the items below each segment's C_MISS0 are shaded here).
template <int BLOCK, int WAVES>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(WAVES > 0 ? WAVES : 1, 8))) void k_wf_shade(
    WfState w, int in_slot, int use_sorted, int split) {
    unsigned long long c0;
    const uint32_t n = chunk_n(w, c0);
    const uint32_t* counts = w.ctrl + C_QCNT + in_slot * kSeg * C_LINE;
    const uint32_t* q = use_sorted ? w.sorted : w.q[in_slot];
    const int out_slot = in_slot ^ 1;
    uint32_t* qout = w.q[out_slot];
    uint32_t* out_counts = w.ctrl + C_QCNT + out_slot * kSeg * C_LINE;
    QueueWalk qw(w.ctrl, C_CUR_SHADE, counts, n, w.grab_shade, nullptr, split ? w.ctrl + C_MISS0 : nullptr);
    // Three items per lane in flight: the one being shaded (A), the next one
    // whose path record and hit are being loaded (B) and the one after it
    // whose queue entry is being loaded (C).  Shading is a chain of
    // dependent gathers at two waves per SIMD; loading B and C during A
    // takes the queue entry -> path record steps off the chain.
    ShadeEvents ev;
    ShadeItem b, c;
    b.has = qw.next(b.i, b.end);
    b.seg = qw.seg;
    if (b.has && b.i < b.end) b.p = q[b.i];
    c.has = b.has && qw.next(c.i, c.end);
    c.seg = qw.seg;
    if (c.has && c.i < c.end) c.p = q[c.i];
    if (b.has && b.i < b.end) b.fetch(w);
    while (b.has) {
        const ShadeItem a = b;
        b = c;
        if (b.has && b.i < b.end) b.fetch(w);
        c.has = c.has && qw.next(c.i, c.end);
        c.seg = qw.seg;
        if (c.has && c.i < c.end) c.p = q[c.i];
        const int seg = a.seg;
        const bool active = a.i < a.end;
        const uint32_t p = a.p;
        PathState st;
        uint4 r = a.rng;
        bool cont = false;
        DeferNee nee;
        nee.base = seg_base(n, seg);
        nee.sa_count = &w.ctrl[C_SACNT + seg * C_LINE];
        nee.se_count = &w.ctrl[C_SECNT + seg * C_LINE];
        nee.sa_o = w.sa_o;
        nee.sa_d = w.sa_d;
        nee.sa_c = w.sa_c;
        nee.se_o = w.se_o;
        nee.se_d = w.se_d;
        nee.se_c = w.se_c;
        nee.sp_count = &w.ctrl[C_SP];
        nee.sp_capacity = w.sp_capacity;
        nee.sp_o = w.sp_o;
        nee.sp_d = w.sp_d;
        nee.sp_c = w.sp_c;
        if (active) {
            nee.path = p;
            a.unpack(st);
            const JobDesc& job = w.jobs[w.njobs == 1 ? 0u : r.x];
            const RenderParams& rp = job.params;
            Hit h;
            h.t = a.hv.x;
            h.u = a.hv.y;
            h.v = a.hv.z;
            h.tri = __float_as_uint(a.hv.w);
            if (st.depth < rp.max_depth) {  // a job of smaller depth than the batch's stops here
                Sampler rng = sampler_of(rp.sampler, r);
                cont = bounce(job.scene, rp, h, st, rng, nee);
                r = sampler_state(rng, r);
                if (cont) st.depth += 1;
                cont = cont && st.depth < rp.max_depth;
                store_path_added(w, p, st, r, cont);
                if (w.shade_ctr) {
                    ev.items += 1;
                    const bool miss = h.tri == 0xFFFFFFFFu;
                    ev.env_lookups += (miss && job.scene.env != nullptr) ? 1u : 0u;
                    ev.hits += miss ? 0u : 1u;
                    ev.light_samples += nee.light_samples;
                    ev.entries += nee.entries;
                    ev.conts += cont ? 1u : 0u;
                    ev.radiance_adds += is_zero(st.lo) ? 0u : 1u;
                }
            }
        }
        // the continuation goes to the segment its item came from
        const uint32_t ko = seg_base(n, seg) + wave_append(&out_counts[seg * C_LINE], cont);
        if (cont) qout[ko] = p;
    }
    if (w.shade_ctr) ev.flush(w.shade_ctr);
}

// The escapes of the bounce (the material-sorted queue's last bin, from each
// segment's C_MISS0 to its end): the environment's radiance onto the path's
// sum, the path ends.  Half of a bounce's items can be escapes; here they
// take a few registers at full occupancy instead of a slot of the surface
// kernel, whose registers hold a whole bounce.  The arithmetic is bounce()'s
// own (escape()'s steps).  An escape is a chain of gathers (queue entry ->
// path record and sum -> texels and pmf entry); three items per lane are in
// flight, each a step further along it: the one being finished (A), the next
// one whose record and sum are being loaded (B) and the one after it whose
// queue entry is being loaded (C).
struct EscapeNee {
    __device__ __forceinline__ void on_env_miss(const DevScene&, const PathState&, float) {}
};

struct EscapeItem {
    bool has = false;  // the wave got a round of items
    uint32_t i = 0, end = 0, p = 0;
    float4 d, tp, lo;
    uint32_t job = 0;
    __device__ __forceinline__ bool valid() const { return has && i < end; }
    __device__ __forceinline__ void fetch(const WfState& w) {
        const PathRec& R = w.path[p];
        d = R.d;
        tp = R.tp;
        job = R.rng.x;
        // the path's sum: no other kernel of this launch writes it (each path
        // is in the queue once; the shadow kernels run in other launches)
        lo = w.lo[p];
    }
};

// Items per grab of the escape kernel (rows of 64): on the 1080p icosphere
// frame 0.56 ms per launch with 4 rows, 0.57 with 16, 0.67 with 64.
constexpr uint32_t kEscapeGrab = 256;

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_wf_shade_escape(WfState w,
                                                                                                     int in_slot) {
    unsigned long long c0;
    const uint32_t n = chunk_n(w, c0);
    const uint32_t* counts = w.ctrl + C_QCNT + in_slot * kSeg * C_LINE;
    QueueWalk qw(w.ctrl, C_CUR_MISS, counts, n, kEscapeGrab, w.ctrl + C_MISS0, nullptr);
    ShadeEvents ev;
    EscapeItem b, c;
    b.has = qw.next(b.i, b.end);
    if (b.valid()) b.p = w.sorted[b.i];
    c.has = b.has && qw.next(c.i, c.end);
    if (c.valid()) c.p = w.sorted[c.i];
    if (b.valid()) b.fetch(w);
    while (b.has) {
        const EscapeItem a = b;
        b = c;
        if (b.valid()) b.fetch(w);
        c.has = c.has && qw.next(c.i, c.end);
        if (c.valid()) c.p = w.sorted[c.i];
        if (!a.valid()) continue;
        const JobDesc& job = w.jobs[w.njobs == 1 ? 0u : a.job];
        PathState st;
        st.d = f3(a.d);
        st.last_pdf = a.d.w;
        st.tp = f3(a.tp);
        const uint32_t bits = __float_as_uint(a.tp.w);
        st.depth = static_cast<int>(bits & 0xFFFFu);
        st.last_delta = (bits >> 31) != 0;
        st.lo = f3(0.0f);
        if (st.depth >= job.params.max_depth) continue;  // a job of smaller depth than the batch's stops here
        const DevScene& s = job.scene;
        if (s.env != nullptr) {
            const EscapeTaps e = escape_taps(s, st.d);
            const float4 t[4] = {s.env[e.tex.idx[0]], s.env[e.tex.idx[1]], s.env[e.tex.idx[2]], s.env[e.tex.idx[3]]};
            const float pmf = env_pdf_wanted(st) && e.pmf_idx >= 0 ? s.env_pmf[e.pmf_idx] : 0.0f;
            EscapeNee nee;
            env_escape(s, st, e, t, pmf, nee);
        }
        if (!is_zero(st.lo)) w.lo[a.p] = make_float4(a.lo.x + st.lo.x, a.lo.y + st.lo.y, a.lo.z + st.lo.z, 0.0f);
        if (w.shade_ctr) {
            ev.items += 1;
            ev.env_lookups += s.env != nullptr ? 1u : 0u;
            ev.radiance_adds += is_zero(st.lo) ? 0u : 1u;
        }
    }
    if (w.shade_ctr) ev.flush(w.shade_ctr);
}

__device__ __forceinline__ void add_unoccluded(const WfState& w, uint32_t p, float4 c, bool atomic) {
    if (atomic) {
        atomicAdd(&w.lo[p].x, c.x);
        atomicAdd(&w.lo[p].y, c.y);
        atomicAdd(&w.lo[p].z, c.z);
    } else {
        float4 l = w.lo[p];
        w.lo[p] = make_float4(l.x + c.x, l.y + c.y, l.z + c.z, l.w);
    }
}

// Any-hit over a segmented shadow queue (area or environment): each path has
// at most one entry, so the contribution is added without atomics.
template <bool COUNT, int BLOCK, int LDS, bool WIDE>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_wf_shadow(WfState w, int count_slot,
                                                                                                int cur_slot,
                                                     const float4* __restrict__ qo, const float4* __restrict__ qd,
                                                     const float4* __restrict__ qc, DevCounters* ctr) {
    __shared__ uint32_t stack[LDS * BLOCK];
    const auto sk = thread_stack<LDS, BLOCK, stack_total<WIDE>()>(w, stack);
    unsigned long long c0;
    const uint32_t n = chunk_n(w, c0);
    const uint32_t* counts = w.ctrl + count_slot;
    QueueWalk qw(w.ctrl, cur_slot, counts, n, w.grab_shadow);
    uint32_t i, end;
    TraceCounters cnt{0, 0};
    uint32_t ovf = 0;
    unsigned long long rays = 0;
    const bool one_job = w.njobs == 1;  // uniform scene: scalar base addresses
    while (qw.next(i, end)) {
        if (i < end) {
            float4 o = qo[i], d = qd[i];
            const uint32_t p = __float_as_uint(d.w);
            rays++;
            const bool occ = one_job ? trace_any_v<WIDE, COUNT>(w.jobs[0].scene, f3(o), f3(d), o.w, sk, cnt, ovf)
                                     : trace_any_v<WIDE, COUNT>(w.jobs[path_job(w, p)].scene, f3(o), f3(d), o.w, sk,
                                                                cnt, ovf);
            if (!occ) add_unoccluded(w, p, qc[i], false);
        }
    }
    flush<COUNT, true>(ctr, rays, cnt.nodes, cnt.tris, ovf);
}

// Any-hit over the flat punctual-light queue (several entries per path:
// contributions are added with float atomics).
template <bool COUNT, int BLOCK, int LDS, bool WIDE>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_wf_shadow_punctual(
    WfState w, DevCounters* ctr) {
    __shared__ uint32_t stack[LDS * BLOCK];
    const auto sk = thread_stack<LDS, BLOCK, stack_total<WIDE>()>(w, stack);
    const uint32_t n = min(w.ctrl[C_SP], w.sp_capacity);
    QueueWalk qw(w.ctrl, C_CUR_SP, nullptr, n, w.grab_shadow);
    uint32_t i, end;
    TraceCounters cnt{0, 0};
    uint32_t ovf = 0;
    unsigned long long rays = 0;
    const bool one_job = w.njobs == 1;
    while (qw.next(i, end)) {
        if (i < end) {
            float4 o = w.sp_o[i], d = w.sp_d[i];
            const uint32_t p = __float_as_uint(d.w);
            rays++;
            const bool occ = one_job ? trace_any_v<WIDE, COUNT>(w.jobs[0].scene, f3(o), f3(d), o.w, sk, cnt, ovf)
                                     : trace_any_v<WIDE, COUNT>(w.jobs[path_job(w, p)].scene, f3(o), f3(d), o.w, sk,
                                                                cnt, ovf);
            if (!occ) add_unoccluded(w, p, w.sp_c[i], true);
        }
    }
    flush<COUNT, true>(ctr, rays, cnt.nodes, cnt.tris, ovf);
}

// Per global pixel touched by the chunk: add its samples of this chunk to the
// pixel's sum, one at a time in sample order (as the megakernel does).  A
// block whose 256 pixels belong to one job and have all their samples in
// this chunk (the usual case) reads their samples, which are contiguous, in
// slices of kFinSlice per pixel through LDS: the loads are coalesced rather
// than one 16-byte stream per lane; each thread then sums its own pixel's
// slice in order.
constexpr int kFinBlock = 256;
constexpr int kFinSlice = 16;
constexpr int kFinRow = 3 * kFinSlice + 1;  // floats per pixel row in LDS (padded: no bank conflicts)

__global__ __launch_bounds__(kFinBlock) void k_wf_finalize(WfState w, uint32_t gp0, uint32_t gp1) {
    __shared__ float stage[kFinBlock * kFinRow];
    const uint32_t t = threadIdx.x;
    const uint32_t gp = gp0 + blockIdx.x * kFinBlock + t;
    unsigned long long c0;
    const uint32_t n = chunk_n(w, c0);
    const bool in = gp < gp1 && gp < w.total_pixels;
    uint32_t j = 0, spf = 0;
    unsigned long long first = 0, a = 0, b = 0;
    if (in) {
        j = job_of_pixel(w, gp);
        const JobDesc& job = w.jobs[j];
        spf = static_cast<uint32_t>(job.params.spf);
        first = job.path_offset + static_cast<unsigned long long>(gp - job.pixel_offset) * spf;
        a = first > c0 ? first : c0;
        b = (first + spf) < (c0 + n) ? (first + spf) : (c0 + n);
    }
    // block-uniform: every thread has a pixel of thread 0's job, all of whose samples are in the chunk
    __shared__ uint32_t job0;
    if (t == 0) job0 = in ? j : 0xFFFFFFFFu;
    __syncthreads();
    const bool whole = in && j == job0 && a == first && b == first + spf;
    if (__syncthreads_and(whole ? 1 : 0)) {
        const unsigned long long base = first - t * static_cast<unsigned long long>(spf) - c0;  // thread 0's first, chunk-local
        F3 sum = f3(0.0f);
        for (uint32_t s0 = 0; s0 < spf; s0 += kFinSlice) {
            const uint32_t cnt = min(static_cast<uint32_t>(kFinSlice), spf - s0);
            for (uint32_t e = t; e < kFinBlock * cnt; e += kFinBlock) {
                const uint32_t px = e / cnt, sm = e % cnt;
                const unsigned long long idx = base + static_cast<unsigned long long>(px) * spf + s0 + sm;
                const float4 l = w.lo[idx];
                if (w.kept_mode == 1) w.kept[c0 + idx] = l;
                float* row = stage + px * kFinRow + 3 * sm;
                row[0] = l.x;
                row[1] = l.y;
                row[2] = l.z;
            }
            __syncthreads();
            const float* row = stage + t * kFinRow;
            for (uint32_t sm = 0; sm < cnt; ++sm)
                sum += clamp_radiance(f3(row[3 * sm], row[3 * sm + 1], row[3 * sm + 2]));
            __syncthreads();
        }
        w.sum[gp] = make_float4(sum.x, sum.y, sum.z, 0.0f);
        return;
    }
    if (!in || a >= b) return;
    F3 sum = first >= c0 ? f3(0.0f) : f3(w.sum[gp]);
    for (unsigned long long g = a; g < b; ++g) {
        const float4 l = w.lo[g - c0];
        if (w.kept_mode == 1) w.kept[g] = l;
        sum += clamp_radiance(f3(l));
    }
    w.sum[gp] = make_float4(sum.x, sum.y, sum.z, 0.0f);
}

// Batch mean to each job's films (path.cu:228-241).
__global__ void k_wf_film(WfState w, DevCounters* ctr, uint32_t count_paths) {
    const uint32_t gp = blockIdx.x * blockDim.x + threadIdx.x;
    if (gp >= w.total_pixels) return;
    const JobDesc& job = w.jobs[job_of_pixel(w, gp)];
    const RenderParams& rp = job.params;
    const uint32_t px = gp - job.pixel_offset;
    float4 sm = w.sum[gp];
    F3 acc = f3(sm) / static_cast<float>(rp.spf);
    float4 nw = make_float4(acc.x, acc.y, acc.z, 1.0f);
    job.film_new[px] = nw;
    float4 a = job.film_acc[px];
    float inv = 1.0f / static_cast<float>(rp.frame_count + 1u);
    a.x = a.x * (1.0f - inv) + nw.x * inv;
    a.y = a.y * (1.0f - inv) + nw.y * inv;
    a.z = a.z * (1.0f - inv) + nw.z * inv;
    a.w = a.w * (1.0f - inv) + nw.w * inv;
    job.film_acc[px] = a;
    if (count_paths && gp == 0) atomicAdd(&ctr->paths, static_cast<unsigned long long>(w.total_paths));
}

// ------------------------------------------------------- adjoint (replay)
// The wavefront form of the path-replay adjoint.  A chunk first runs the
// primal bounce loop (k_wf_shade: each path's radiance L ends in lo);
// k_wfa_seed keeps L and the path's adjoint g = d loss / d L; the chunk's
// paths are then generated again from the same sampler state and replayed
// by k_wfa_shade and the k_wfa_shadow kernels:
//   * the replay shade appends NEE shadow rays with their sensitivities
//     (d/dtheta of T Le f w / p at the light direction and, for environment
//     samples, the weight of d/dLe); a shadow kernel applies them to the
//     parameter buffers when the ray is unoccluded;
//   * a miss applies the gradient of T env w to the environment texels;
//   * a sampled continuation stores rho_c = g_c (df_c/dtheta) / f_c for the
//     path; the next vertex's shade, when everything the path gathered up to
//     the previous vertex is in lo (its shadow rays have been resolved),
//     applies (L - lo) . rho: the radiance still to come times the relative
//     change of the throughput.  A path that ends at a vertex (Russian
//     roulette, maximum depth) has nothing still to come.
// Pdfs, MIS weights and Russian-roulette factors are detached, as in the
// megakernel adjoint.  All jobs are views of one scene.

constexpr int kPend = 5;     // float4 per path: rho[3][5], material | kNoVertex, tu, tv
constexpr int kPayload = 3;  // float4 per shadow entry: gv[5], material, tu, tv, env weight rgb
constexpr uint32_t kNoVertex = 0xFFFFFFFFu;
constexpr uint32_t kWfLdsMaterials = 64;  // material sensitivities summed in LDS up to this many materials
constexpr int kReplayStackLds = 16;        // replay shadow kernels: traversal stack entries in LDS (rest spilled)

__device__ __forceinline__ float wave_sum_f(float v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ void atomic_add3(float* dst, F3 v) {
    if (v.x != 0.0f) atomicAdd(dst + 0, v.x);
    if (v.y != 0.0f) atomicAdd(dst + 1, v.y);
    if (v.z != 0.0f) atomicAdd(dst + 2, v.z);
}

// dst[4 idx + 0..2] += v on every lane with `valid`.  Lanes of the wave that
// add to the same texel are summed first and add with one atomic (a pixel's
// samples are neighbours in the queues and mostly touch the same texels).
// Every lane of the wave calls it.
__device__ void wave_add_texel(float* dst, uint32_t idx, F3 v, bool valid) {
    unsigned long long pending = __ballot(valid);
    const int lane = lane_id();
    for (int round = 0; pending != 0ull && round < 8; ++round) {
        const int leader = __ffsll(static_cast<long long>(pending)) - 1;
        const uint32_t lidx = __shfl(idx, leader, 64);
        const bool mine = ((pending >> lane) & 1ull) != 0ull && idx == lidx;
        const unsigned long long m = __ballot(mine);
        pending &= ~m;
        if (__popcll(m) == 1) {
            if (lane == leader) atomic_add3(dst + 4 * static_cast<size_t>(idx), v);
            continue;
        }
        const F3 sum = f3(wave_sum_f(mine ? v.x : 0.0f), wave_sum_f(mine ? v.y : 0.0f), wave_sum_f(mine ? v.z : 0.0f));
        if (lane == leader) atomic_add3(dst + 4 * static_cast<size_t>(lidx), sum);
    }
    if ((pending >> lane) & 1ull) atomic_add3(dst + 4 * static_cast<size_t>(idx), v);
}

__device__ __forceinline__ void add_material(float* macc, const AdjointGrads& G, uint32_t mi, const float m[kDP]) {
    if (mi < kWfLdsMaterials) {
#pragma unroll
        for (int k = 0; k < kDP; ++k)
            if (m[k] != 0.0f) atomicAdd(&macc[mi * kDP + k], m[k]);
        return;
    }
    if (G.base)
        for (int k = 0; k < 3; ++k)
            if (m[k] != 0.0f) atomicAdd(&G.base[3 * mi + k], m[k]);
    if (G.rough && m[kDRough] != 0.0f) atomicAdd(&G.rough[mi], m[kDRough]);
    if (G.metal && m[kDMetal] != 0.0f) atomicAdd(&G.metal[mi], m[kDMetal]);
}

__device__ __forceinline__ void clear_materials(float* macc) {
    for (uint32_t k = threadIdx.x; k < kWfLdsMaterials * kDP; k += blockDim.x) macc[k] = 0.0f;
    __syncthreads();
}

// The block's LDS material sums to the parameter buffers (kernel end).
__device__ __forceinline__ void flush_materials(const float* macc, const AdjointGrads& G, uint32_t nmat) {
    __syncthreads();
    const uint32_t n = min(nmat, kWfLdsMaterials) * kDP;
    for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) {
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

// A vertex's sensitivities gv (lanes with `valid`) through the parameter
// chain to the buffers.  Every lane of the wave calls it.
__device__ void apply_vertex(const DevScene& s, const AdjointGrads& G, float* macc, bool valid, uint32_t mi,
                             float tu, float tv, const float gv[kDP]) {
    bool any = false;
#pragma unroll
    for (int k = 0; k < kDP; ++k) any = any || gv[k] != 0.0f;
    valid = valid && any;
    ParamChain pc{};
    if (valid) {
        param_chain(s, s.materials[mi], tu, tv, gv, pc);
        add_material(macc, G, mi, pc.m);
    }
    const bool tex = valid && pc.tex && G.tex != nullptr;
    if (__ballot(tex) == 0ull) return;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        wave_add_texel(G.tex, tex ? pc.tidx[k] : 0u, tex ? pc.k3 * pc.tw[k] : f3(0.0f), tex && pc.tw[k] != 0.0f);
}

// d/d(env texels) of coef . env(d) (lanes with `valid`).  Every lane of the wave calls it.
__device__ void apply_env(const DevScene& s, const AdjointGrads& G, bool valid, F3 d, F3 coef) {
    valid = valid && G.env != nullptr && s.env != nullptr && !is_zero(coef);
    if (__ballot(valid) == 0ull) return;
    TexTaps t{};
    if (valid) t = env_taps(s, d);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const bool on = valid && t.w[k] != 0.0f;
        wave_add_texel(G.env, on ? t.idx[k] : 0u, on ? coef * (t.w[k] * kTwoPi) : f3(0.0f), on);
    }
}

// The splat stage's form of apply_env: lanes with `valid` record the lookup's
// (u, v) and their coefficient in their segment's part of the records, keyed
// by the splat tile of the lookup's first tap.  Every lane of the wave calls
// it.
__device__ void env_record(const WfState& w, const DevScene& s, const AdjointGrads& G, bool valid, F3 d, F3 coef,
                           int seg, uint32_t n) {
    valid = valid && G.env != nullptr && s.env != nullptr && !is_zero(coef);
    float u = 0.0f, v = 0.0f;
    TexTaps t{};
    if (valid) {
        dir_to_uv(d, u, v);  // env_taps()'s steps: the splat repeats tex_taps() on (u, v)
        t = tex_taps(s.env_w, s.env_h, s.env_wrap_s, s.env_wrap_t, u, v);
    }
    const uint32_t k = seg_base(n, seg) + wave_append(&w.ctrl[C_ESCNT + seg * C_LINE], valid);
    if (!valid) return;
    w.esc_rec[k] = make_float4(u, v, coef.x, coef.y);
    w.esc_cz[k] = coef.z;
    const uint32_t ew = static_cast<uint32_t>(s.env_w);
    const uint32_t x0 = t.idx[0] % ew, y0 = t.idx[0] / ew;
    w.esc_key[k] = min((y0 / kSplatTileH) * w.env_tiles_x + x0 / kSplatTileW, w.env_tiles - 1);
}

// The counting sort's scatter for the environment-gradient records: moves
// each record to its place in tile order (the splat then reads them in order).
__global__ __launch_bounds__(256) void k_wfa_env_scatter(WfState w, SortSpec a, uint32_t nbins) {
    extern __shared__ uint32_t off[];
    const int seg = static_cast<int>(blockIdx.x % kSeg), part = static_cast<int>(blockIdx.x / kSeg);
    const uint32_t* src = a.hist + (static_cast<size_t>(seg) * a.parts + part) * nbins;
    for (uint32_t b = threadIdx.x; b < nbins; b += blockDim.x) off[b] = src[b];
    __syncthreads();
    unsigned long long c0;
    const uint32_t n = chunk_n(w, c0);
    uint32_t b0, b1;
    sort_part_range(n, a.counts, seg, part, a.parts, b0, b1);
    for (uint32_t i = b0 + threadIdx.x; i < b1; i += blockDim.x) {
        const uint32_t o = atomicAdd(&off[a.key[i]], 1u);
        w.esc_srec[o] = w.esc_rec[i];
        w.esc_scz[o] = w.esc_cz[i];
    }
}

// The splat of one (segment, tile), in kSplatParts blocks: block blockIdx.x
// = (tile * kSeg + segment) * kSplatParts + part sums its part of the tile's
// records (a contiguous run of the segment's sorted records) in LDS over the
// tile and its one-texel apron on the right and bottom (a first tap in the
// tile has its other taps there, except where the lookup wraps; such taps
// are added directly), then adds each nonzero texel channel to the
// environment gradient.
constexpr int kSplatW = kSplatTileW + 1, kSplatH = kSplatTileH + 1;
__global__ __launch_bounds__(256) void k_wfa_env_splat(WfState w) {
    __shared__ float acc[kSplatW * kSplatH * 3];
    const uint32_t part = blockIdx.x % kSplatParts, st = blockIdx.x / kSplatParts;
    const uint32_t seg = st % kSeg, tile = st / kSeg;
    const uint32_t ntiles = w.env_tiles;
    unsigned long long c0;
    const uint32_t n = chunk_n(w, c0);
    const uint32_t* h = w.esc_hist + static_cast<size_t>(seg) * kSplatSortParts * ntiles;  // part 0: the tiles' starts
    const uint32_t t0 = h[tile];
    const uint32_t t1 =
        tile + 1 < ntiles ? h[tile + 1] : seg_base(n, static_cast<int>(seg)) + w.ctrl[C_ESCNT + seg * C_LINE];
    if (t0 >= t1) return;  // the whole block
    const uint32_t len = t1 - t0;
    const uint32_t r0 = t0 + static_cast<uint32_t>((static_cast<unsigned long long>(len) * part) / kSplatParts);
    const uint32_t r1 = t0 + static_cast<uint32_t>((static_cast<unsigned long long>(len) * (part + 1)) / kSplatParts);
    if (r0 >= r1) return;  // the whole block
    const AdjointGrads G = *w.grads;
    const DevScene& s = w.jobs[0].scene;
    const uint32_t ew = static_cast<uint32_t>(s.env_w), eh = static_cast<uint32_t>(s.env_h);
    const uint32_t tx0 = (tile % w.env_tiles_x) * kSplatTileW, ty0 = (tile / w.env_tiles_x) * kSplatTileH;
    for (int j = threadIdx.x; j < kSplatW * kSplatH * 3; j += blockDim.x) acc[j] = 0.0f;
    __syncthreads();
    for (uint32_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
        const float4 a = w.esc_srec[r];
        const F3 coef = f3(a.z, a.w, w.esc_scz[r]);
        const TexTaps t = tex_taps(s.env_w, s.env_h, s.env_wrap_s, s.env_wrap_t, a.x, a.y);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (t.w[k] == 0.0f) continue;
            const F3 v = coef * (t.w[k] * kTwoPi);
            const uint32_t x = t.idx[k] % ew, y = t.idx[k] / ew;
            if (x >= tx0 && x - tx0 < static_cast<uint32_t>(kSplatW) && y >= ty0 &&
                y - ty0 < static_cast<uint32_t>(kSplatH)) {
                float* dst = acc + 3 * ((y - ty0) * kSplatW + (x - tx0));
                if (v.x != 0.0f) atomicAdd(dst + 0, v.x);
                if (v.y != 0.0f) atomicAdd(dst + 1, v.y);
                if (v.z != 0.0f) atomicAdd(dst + 2, v.z);
            } else {
                atomic_add3(G.env + 4 * static_cast<size_t>(t.idx[k]), v);
            }
        }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < kSplatW * kSplatH * 3; j += blockDim.x) {
        const float v = acc[j];
        if (v == 0.0f) continue;
        const uint32_t cell = static_cast<uint32_t>(j) / 3, ch = static_cast<uint32_t>(j) % 3;
        const uint32_t x = tx0 + cell % kSplatW, y = ty0 + cell / kSplatW;
        if (x < ew && y < eh) atomicAdd(G.env + 4 * (static_cast<size_t>(y) * ew + x) + ch, v);
    }
}

__global__ void k_wfa_seed(WfState w) {
    unsigned long long c0;
    const uint32_t n = chunk_n(w, c0);
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const unsigned long long g = c0 + p;
    const JobDesc& job = w.jobs[job_of_path(w, g)];
    const uint32_t spf = static_cast<uint32_t>(job.params.spf);
    const uint32_t pixel = static_cast<uint32_t>((g - job.path_offset) / spf);
    const float4 L = w.kept_mode == 2 ? w.kept[g] : w.lo[p];
    const float4 gp = job.img_grad[pixel];
    const float inv = 1.0f / static_cast<float>(spf);
    w.adj_L[p] = L;
    // fminf(L, 100) passes the gradient below the clamp (path.cu:201-204)
    w.adj_g[p] = make_float4(L.x <= 100.0f ? gp.x * inv : 0.0f, L.y <= 100.0f ? gp.y * inv : 0.0f,
                             L.z <= 100.0f ? gp.z * inv : 0.0f, 0.0f);
}

// Visibility policy of the replay: DeferNee's queues, each entry with its
// sensitivities; the continuation's rho goes to the path's pending record.
// NP: the partials the BSDF derivatives carry (to_dual): 3 when no
// roughness or metallic gradient is wanted, the other two then stay zero.
template <int NP>
struct AdjDeferNee {
    static constexpr bool kDrawAhead = false;
    uint32_t path;
    uint32_t base;
    uint32_t *sa_count, *se_count;
    float4 *sa_o, *sa_d, *sa_c, *sa_x, *se_o, *se_d, *se_c, *se_x;
    uint32_t* sp_count;
    uint32_t sp_capacity;
    float4 *sp_o, *sp_d, *sp_c, *sp_x;
    float4* pend;  // this path's kPend records
    F3 g;
    bool wrote_pend;
    bool miss;
    F3 miss_coef;

    __device__ __forceinline__ static void put_payload(float4* x, const float gv[kDP], uint32_t mi, const ShadeCtx& c,
                                                       F3 ec) {
        x[0] = make_float4(gv[0], gv[1], gv[2], gv[3]);
        x[1] = make_float4(gv[4], __uint_as_float(mi), c.tu, c.tv);
        x[2] = make_float4(ec.x, ec.y, ec.z, 0.0f);
    }

    __device__ void direct(const DevScene& s, PathState& st, bool a, const ShadowRay& ra, F3 xa, bool e,
                           const ShadowRay& re, F3 xe, const LightSample& la, const LightSample& le,
                           const ShadeCtx& x) {
        const uint32_t mi = static_cast<uint32_t>(x.mat - s.materials);
        const uint32_t ka = base + wave_append(sa_count, a);
        if (a) {
            sa_o[ka] = make_float4(ra.o.x, ra.o.y, ra.o.z, ra.tmax);
            sa_d[ka] = make_float4(ra.d.x, ra.d.y, ra.d.z, __uint_as_float(path));
            sa_c[ka] = make_float4(xa.x, xa.y, xa.z, 0.0f);
            float gv[kDP] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            F3 ec = f3(0.0f);
            light_grad<NP>(st, g, la, x, gv, ec);
            put_payload(sa_x + kPayload * static_cast<size_t>(ka), gv, mi, x, f3(0.0f));
        }
        const uint32_t ke = base + wave_append(se_count, e);
        if (e) {
            se_o[ke] = make_float4(re.o.x, re.o.y, re.o.z, re.tmax);
            se_d[ke] = make_float4(re.d.x, re.d.y, re.d.z, __uint_as_float(path));
            se_c[ke] = make_float4(xe.x, xe.y, xe.z, 0.0f);
            float gv[kDP] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            F3 ec = f3(0.0f);
            if (!light_grad<NP>(st, g, le, x, gv, ec)) ec = f3(0.0f);
            put_payload(se_x + kPayload * static_cast<size_t>(ke), gv, mi, x, ec);
        }
    }
    __device__ __forceinline__ void on_env_miss(const DevScene&, const PathState& st, float mis) {
        miss = true;
        miss_coef = f3(g.x * st.tp.x * mis, g.y * st.tp.y * mis, g.z * st.tp.z * mis);
    }
    __device__ void on_bsdf_sample(const DevScene& s, const PathState&, const BsdfSample& bs, const ShadeCtx& x) {
        const SV3<Dual<NP>> f = continuation_f<NP>(x, bs);
        const float rx = f.x.v != 0.0f ? g.x / f.x.v : 0.0f;
        const float ry = f.y.v != 0.0f ? g.y / f.y.v : 0.0f;
        const float rz = f.z.v != 0.0f ? g.z / f.z.v : 0.0f;
        float r[16];
#pragma unroll
        for (int j = 0; j < kDP; ++j) {
            r[j] = j < NP ? rx * f.x.d[j] : 0.0f;
            r[kDP + j] = j < NP ? ry * f.y.d[j] : 0.0f;
            r[2 * kDP + j] = j < NP ? rz * f.z.d[j] : 0.0f;
        }
        r[15] = __uint_as_float(static_cast<uint32_t>(x.mat - s.materials));
        pend[0] = make_float4(r[0], r[1], r[2], r[3]);
        pend[1] = make_float4(r[4], r[5], r[6], r[7]);
        pend[2] = make_float4(r[8], r[9], r[10], r[11]);
        pend[3] = make_float4(r[12], r[13], r[14], r[15]);
        pend[4] = make_float4(x.tu, x.tv, 0.0f, 0.0f);
        wrote_pend = true;
    }
    __device__ void punctual_lights(const DevScene& s, PathState& st, const ShadeCtx& x) {
        const uint32_t mi = static_cast<uint32_t>(x.mat - s.materials);
        for (uint32_t i = 0; i < s.num_lights; ++i) {
            F3 wi, inten;
            float dist;
            if (!punctual_light_sample(s, i, x.p, x.n, wi, inten, dist)) continue;
            BsdfEvalT<float> b = eval_bsdf(*x.me, x.wo, wi, x.n, true);
            F3 c = st.tp * (inten * f3(b.f.x, b.f.y, b.f.z));
            if (is_zero(c)) continue;
            uint32_t k = atomicAdd(sp_count, 1u);
            if (k >= sp_capacity) continue;  // cannot happen: capacity = paths x lights
            ShadowRay r = make_shadow(x.p, x.n, wi, dist);
            sp_o[k] = make_float4(r.o.x, r.o.y, r.o.z, r.tmax);
            sp_d[k] = make_float4(r.d.x, r.d.y, r.d.z, __uint_as_float(path));
            sp_c[k] = make_float4(c.x, c.y, c.z, 0.0f);
            float gv[kDP] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            punctual_grad<NP>(st, g, wi, inten, x, gv);
            put_payload(sp_x + kPayload * static_cast<size_t>(k), gv, mi, x, f3(0.0f));
        }
    }
};

// Built for 2 waves per SIMD (registers + accumulation registers within 256).
// The previous vertex's pending continuation of path p at the start of a
// bounce: its sensitivities (L - lo) . rho, material and texture coordinates
// (pv false when the path has none).
struct PendingVertex {
    bool pv = false;
    uint32_t mi = 0;
    float tu = 0.0f, tv = 0.0f;
    float gv[kDP] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
};
__device__ __forceinline__ void load_pending(const WfState& w, uint32_t p, const PathState& st, PendingVertex& v) {
    if (st.depth <= 0) return;
    const float4* pend = w.pend + kPend * static_cast<size_t>(p);
    const float4 p3 = pend[3];
    const uint32_t m = __float_as_uint(p3.w);
    if (m == kNoVertex) return;
    const float4 p0 = pend[0], p1 = pend[1], p2 = pend[2], p4 = pend[4];
    const float rho[15] = {p0.x, p0.y, p0.z, p0.w, p1.x, p1.y, p1.z, p1.w, p2.x, p2.y, p2.z, p2.w, p3.x, p3.y, p3.z};
    const F3 rem = f3(w.adj_L[p]) - st.lo;
#pragma unroll
    for (int j = 0; j < kDP; ++j) v.gv[j] = rem.x * rho[j] + rem.y * rho[kDP + j] + rem.z * rho[2 * kDP + j];
    v.pv = true;
    v.mi = m;
    v.tu = p4.x;
    v.tv = p4.y;
}

// Built for 2 waves per SIMD with five partials (registers + accumulation
// registers within 256), for 3 with the base colour's three (168 registers,
// the rest spilled: faster than 2 waves, slower than 4).
// split: the escaped paths (each segment's items from its C_MISS0 on) are
// left to k_wfa_shade_escape.
template <int BLOCK, int NP>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(NP == 3 ? 3 : 2, 8))) void k_wfa_shade(WfState w,
                                                                                                 int in_slot,
                                                                                                 int use_sorted,
                                                                                                 int split) {
    __shared__ float macc[kWfLdsMaterials * kDP];
    clear_materials(macc);
    const AdjointGrads G = *w.grads;
    const DevScene& s0 = w.jobs[0].scene;
    unsigned long long c0;
    const uint32_t n = chunk_n(w, c0);
    const uint32_t* counts = w.ctrl + C_QCNT + in_slot * kSeg * C_LINE;
    const uint32_t* q = use_sorted ? w.sorted : w.q[in_slot];
    const int out_slot = in_slot ^ 1;
    uint32_t* qout = w.q[out_slot];
    uint32_t* out_counts = w.ctrl + C_QCNT + out_slot * kSeg * C_LINE;
    QueueWalk qw(w.ctrl, C_CUR_SHADE, counts, n, w.grab_shade, nullptr, split ? w.ctrl + C_MISS0 : nullptr);
    uint32_t i, end;
    while (qw.next(i, end)) {
        const int seg = qw.seg;
        const bool active = i < end;
        uint32_t p = 0;
        PathState st{};
        uint4 r = make_uint4(0, 0, 0, 0);
        bool cont = false;
        AdjDeferNee<NP> nee;
        nee.base = seg_base(n, seg);
        nee.sa_count = &w.ctrl[C_SACNT + seg * C_LINE];
        nee.se_count = &w.ctrl[C_SECNT + seg * C_LINE];
        nee.sa_o = w.sa_o;
        nee.sa_d = w.sa_d;
        nee.sa_c = w.sa_c;
        nee.sa_x = w.sa_x;
        nee.se_o = w.se_o;
        nee.se_d = w.se_d;
        nee.se_c = w.se_c;
        nee.se_x = w.se_x;
        nee.sp_count = &w.ctrl[C_SP];
        nee.sp_capacity = w.sp_capacity;
        nee.sp_o = w.sp_o;
        nee.sp_d = w.sp_d;
        nee.sp_c = w.sp_c;
        nee.sp_x = w.sp_x;
        nee.wrote_pend = false;
        nee.miss = false;
        nee.miss_coef = f3(0.0f);
        // the previous vertex's continuation, applied with (L - lo)
        PendingVertex pv;
        if (active) {
            p = q[i];
            nee.path = p;
            nee.pend = w.pend + kPend * static_cast<size_t>(p);
            load_path(w, p, st, r);
            nee.g = f3(w.adj_g[p]);
            load_pending(w, p, st, pv);
            const JobDesc& job = w.jobs[w.njobs == 1 ? 0u : r.x];
            const RenderParams& rp = job.params;
            float4 hv = w.hit[p];
            Hit h;
            h.t = hv.x;
            h.u = hv.y;
            h.v = hv.z;
            h.tri = __float_as_uint(hv.w);
            if (st.depth < rp.max_depth) {
                Sampler rng = sampler_of(rp.sampler, r);
                cont = bounce(job.scene, rp, h, st, rng, nee);
                r = sampler_state(rng, r);
                if (cont) st.depth += 1;
                cont = cont && st.depth < rp.max_depth;
                store_path(w, p, st, r);
                if (cont && !nee.wrote_pend) nee.pend[3] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(kNoVertex));
            }
        }
        apply_vertex(s0, G, macc, pv.pv, pv.mi, pv.tu, pv.tv, pv.gv);
        apply_env(s0, G, nee.miss, st.d, nee.miss_coef);
        const uint32_t ko = seg_base(n, seg) + wave_append(&out_counts[seg * C_LINE], cont);
        if (cont) qout[ko] = p;
    }
    flush_materials(macc, G, s0.num_materials);
}

// The replay's escaped paths (the material-sorted queue's last bin): the
// previous vertex's pending continuation, the environment's radiance onto the
// path's sum and its gradient to the environment texels; the path ends.  The
// same steps as k_wfa_shade takes for a miss, in a kernel light enough for
// several waves per SIMD.
struct MissNee {
    F3 g;
    bool miss = false;
    F3 coef = f3(0.0f);
    __device__ __forceinline__ void on_env_miss(const DevScene&, const PathState& st, float mis) {
        miss = true;
        coef = f3(g.x * st.tp.x * mis, g.y * st.tp.y * mis, g.z * st.tp.z * mis);
    }
};

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_wfa_shade_escape(WfState w,
                                                                                                        int in_slot) {
    __shared__ float macc[kWfLdsMaterials * kDP];
    clear_materials(macc);
    const AdjointGrads G = *w.grads;
    const DevScene& s0 = w.jobs[0].scene;
    unsigned long long c0;
    const uint32_t n = chunk_n(w, c0);
    const uint32_t* counts = w.ctrl + C_QCNT + in_slot * kSeg * C_LINE;
    QueueWalk qw(w.ctrl, C_CUR_MISS, counts, n, kEscapeGrab, w.ctrl + C_MISS0, nullptr);
    const bool splat = w.env_tiles != 0;
    uint32_t i, end;
    while (qw.next(i, end)) {
        const int seg = qw.seg;
        PendingVertex pv;
        MissNee nee;
        F3 dir = f3(0.0f);
        bool first = false;  // an escape of the camera ray
        if (i < end) {
            const uint32_t p = w.sorted[i];
            PathState st;
            uint4 r;
            load_path(w, p, st, r);
            first = st.depth == 0;
            nee.g = f3(w.adj_g[p]);
            load_pending(w, p, st, pv);
            const JobDesc& job = w.jobs[w.njobs == 1 ? 0u : r.x];
            if (st.depth < job.params.max_depth) {
                escape(job.scene, st, nee);
                w.lo[p] = make_float4(st.lo.x, st.lo.y, st.lo.z, 0.0f);
                dir = st.d;
            }
        }
        apply_vertex(s0, G, macc, pv.pv, pv.mi, pv.tu, pv.tv, pv.gv);
        // camera rays that escape are neighbours in the queue and mostly share
        // texels (merged within the wave); later escapes go to the splat
        apply_env(s0, G, nee.miss && (first || !splat), dir, nee.coef);
        if (splat) env_record(w, s0, G, nee.miss && !first, dir, nee.coef, seg, n);
    }
    flush_materials(macc, G, s0.num_materials);
}

// Any-hit over a segmented shadow queue of the replay: an unoccluded ray
// adds its contribution and applies its sensitivities.
template <bool COUNT, int BLOCK, bool ENV, bool WIDE>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_wfa_shadow(WfState w, int count_slot, int cur_slot,
                                                      const float4* __restrict__ qo, const float4* __restrict__ qd,
                                                      const float4* __restrict__ qc, const float4* __restrict__ qx,
                                                      DevCounters* ctr) {
    __shared__ uint32_t stack[kReplayStackLds * BLOCK];
    __shared__ float macc[kWfLdsMaterials * kDP];
    clear_materials(macc);
    const auto sk = thread_stack<kReplayStackLds, BLOCK, stack_total<WIDE>()>(w, stack);
    const AdjointGrads G = *w.grads;
    const DevScene& s0 = w.jobs[0].scene;
    unsigned long long c0;
    const uint32_t n = chunk_n(w, c0);
    const uint32_t* counts = w.ctrl + count_slot;
    QueueWalk qw(w.ctrl, cur_slot, counts, n, w.grab_shadow);
    uint32_t i, end;
    TraceCounters cnt{0, 0};
    uint32_t ovf = 0;
    unsigned long long rays = 0;
    while (qw.next(i, end)) {
        bool vis = false;
        uint32_t mi = 0;
        float tu = 0.0f, tv = 0.0f;
        float gv[kDP] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        F3 ec = f3(0.0f), dir = f3(0.0f);
        if (i < end) {
            float4 o = qo[i], d = qd[i];
            const uint32_t p = __float_as_uint(d.w);
            rays++;
            if (!trace_any_v<WIDE, COUNT>(s0, f3(o), f3(d), o.w, sk, cnt, ovf)) {
                add_unoccluded(w, p, qc[i], false);
                const float4 x0 = qx[kPayload * static_cast<size_t>(i)];
                const float4 x1 = qx[kPayload * static_cast<size_t>(i) + 1];
                gv[0] = x0.x;
                gv[1] = x0.y;
                gv[2] = x0.z;
                gv[3] = x0.w;
                gv[4] = x1.x;
                mi = __float_as_uint(x1.y);
                tu = x1.z;
                tv = x1.w;
                if (ENV) ec = f3(qx[kPayload * static_cast<size_t>(i) + 2]);
                dir = f3(d);
                vis = true;
            }
        }
        apply_vertex(s0, G, macc, vis, mi, tu, tv, gv);
        if (ENV && w.env_tiles) env_record(w, s0, G, vis, dir, ec, qw.seg, n);
        else if (ENV) apply_env(s0, G, vis, dir, ec);
    }
    flush<COUNT, true>(ctr, rays, cnt.nodes, cnt.tris, ovf);
    flush_materials(macc, G, s0.num_materials);
}

template <bool COUNT, int BLOCK, bool WIDE>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_wfa_shadow_punctual(WfState w, DevCounters* ctr) {
    __shared__ uint32_t stack[kReplayStackLds * BLOCK];
    __shared__ float macc[kWfLdsMaterials * kDP];
    clear_materials(macc);
    const auto sk = thread_stack<kReplayStackLds, BLOCK, stack_total<WIDE>()>(w, stack);
    const AdjointGrads G = *w.grads;
    const DevScene& s0 = w.jobs[0].scene;
    const uint32_t n = min(w.ctrl[C_SP], w.sp_capacity);
    QueueWalk qw(w.ctrl, C_CUR_SP, nullptr, n, w.grab_shadow);
    uint32_t i, end;
    TraceCounters cnt{0, 0};
    uint32_t ovf = 0;
    unsigned long long rays = 0;
    while (qw.next(i, end)) {
        bool vis = false;
        uint32_t mi = 0;
        float tu = 0.0f, tv = 0.0f;
        float gv[kDP] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        if (i < end) {
            float4 o = w.sp_o[i], d = w.sp_d[i];
            const uint32_t p = __float_as_uint(d.w);
            rays++;
            if (!trace_any_v<WIDE, COUNT>(s0, f3(o), f3(d), o.w, sk, cnt, ovf)) {
                add_unoccluded(w, p, w.sp_c[i], true);
                const float4 x0 = w.sp_x[kPayload * static_cast<size_t>(i)];
                const float4 x1 = w.sp_x[kPayload * static_cast<size_t>(i) + 1];
                gv[0] = x0.x;
                gv[1] = x0.y;
                gv[2] = x0.z;
                gv[3] = x0.w;
                gv[4] = x1.x;
                mi = __float_as_uint(x1.y);
                tu = x1.z;
                tv = x1.w;
                vis = true;
            }
        }
        apply_vertex(s0, G, macc, vis, mi, tu, tv, gv);
    }
    flush<COUNT, true>(ctr, rays, cnt.nodes, cnt.tris, ovf);
    flush_materials(macc, G, s0.num_materials);
}

// Batch mean of the replay pass to each job's replay image (when given).
__global__ void k_wfa_replay_out(WfState w) {
    const uint32_t gp = blockIdx.x * blockDim.x + threadIdx.x;
    if (gp >= w.total_pixels) return;
    const JobDesc& job = w.jobs[job_of_pixel(w, gp)];
    if (job.replay == nullptr) return;
    const F3 acc = f3(w.sum[gp]) / static_cast<float>(job.params.spf);
    job.replay[gp - job.pixel_offset] = make_float4(acc.x, acc.y, acc.z, 1.0f);
}

template <class T>
bool dalloc(T*& p, size_t n) {
    if (p) (void)hipFree(p);
    p = nullptr;
    return hipMalloc(reinterpret_cast<void**>(&p), sizeof(T) * std::max<size_t>(n, 1)) == hipSuccess;
}

}  // namespace

// ------------------------------------------------------------------ host side

// Event pairs around launches, summed per stage after the batch.
struct StageTimer {
    hipStream_t stream = nullptr;
    hipEvent_t open = nullptr;
    int open_stage = 0;
    std::vector<int> stages;
    std::vector<hipEvent_t> begins, ends;
    void begin(int stage) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return;
        (void)hipEventRecord(e, stream);
        open = e;
        open_stage = stage;
    }
    void end() {
        if (!open) return;
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return;
        (void)hipEventRecord(e, stream);
        stages.push_back(open_stage);
        begins.push_back(open);
        ends.push_back(e);
        open = nullptr;
    }
    void collect(float* ms, uint32_t* n) {
        for (int k = 0; k < WF_STAGES; ++k) {
            ms[k] = 0.0f;
            n[k] = 0;
        }
        for (size_t i = 0; i < stages.size(); ++i) {
            float t = 0.0f;
            (void)hipEventElapsedTime(&t, begins[i], ends[i]);
            ms[stages[i]] += t;
            n[stages[i]] += 1;
            (void)hipEventDestroy(begins[i]);
            (void)hipEventDestroy(ends[i]);
        }
        stages.clear();
        begins.clear();
        ends.clear();
    }
};

struct Timed {
    StageTimer* t;
    Timed(StageTimer* timer, int stage) : t(timer) {
        if (t) t->begin(stage);
    }
    ~Timed() {
        if (t) t->end();
    }
};

class Wavefront {
public:
    WfState st{};
    float stage_ms[WF_STAGES] = {};
    uint32_t stage_launches[WF_STAGES] = {};
    JobDesc* d_jobs = nullptr;
    uint32_t job_capacity = 0;
    size_t spill_capacity = 0;
    uint32_t capacity = 0, sp_capacity = 0, pixels = 0, nbins = 0;
    uint32_t adj_capacity = 0, adj_sp_capacity = 0;
    AdjointGrads* d_grads = nullptr;
    int num_cus = 0;
    // graph cache: [0] the forward render, [1] the adjoint (alternating
    // forward and backward batches keep both)
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    hipGraph_t graph[2] = {nullptr, nullptr};
    uint64_t key[2][14] = {};
    // kept per-path radiance (WfState::kept) and what it was rendered from:
    // the caller's key, then each job's scene view and parameters
    float4* kept = nullptr;
    unsigned long long kept_capacity = 0;
    std::vector<unsigned char> kept_sig;
    bool last_reused = false;

    ~Wavefront() { release(); }

    void release_graph(int k) {
        if (exec[k]) (void)hipGraphExecDestroy(exec[k]);
        if (graph[k]) (void)hipGraphDestroy(graph[k]);
        exec[k] = nullptr;
        graph[k] = nullptr;
    }
    void release_graph() {
        release_graph(0);
        release_graph(1);
    }
    void release_kept() {
        if (kept) (void)hipFree(kept);
        kept = nullptr;
        kept_capacity = 0;
        kept_sig.clear();
    }
    void release() {
        release_kept();
        release_graph();
        void* ps[] = {st.path, st.lo, st.hit, st.q[0], st.q[1], st.sorted, st.qkey, st.sa_o, st.sa_d,
                      st.sa_c, st.se_o, st.se_d, st.se_c, st.sp_o, st.sp_d, st.sp_c, st.blockhist, st.sum, st.ctrl,
                      d_jobs, st.spill, st.adj_L, st.adj_g, st.pend, st.sa_x, st.se_x, st.sp_x, st.esc_rec,
                      st.esc_cz, st.esc_key, st.esc_srec, st.esc_scz, st.esc_hist, d_grads};
        for (void* p : ps)
            if (p) (void)hipFree(p);
        st = WfState{};
        d_jobs = nullptr;
        d_grads = nullptr;
        job_capacity = capacity = sp_capacity = pixels = nbins = 0;
        adj_capacity = adj_sp_capacity = 0;
        spill_capacity = 0;
    }

    // Replay state of the adjoint for npaths paths in flight (after ensure()).
    bool ensure_adjoint(uint32_t npaths) {
        bool ok = true;
        if (npaths > adj_capacity) {
            ok = ok && dalloc(st.adj_L, npaths) && dalloc(st.adj_g, npaths) &&
                 dalloc(st.pend, static_cast<size_t>(kPend) * npaths) &&
                 dalloc(st.sa_x, static_cast<size_t>(kPayload) * npaths) &&
                 dalloc(st.se_x, static_cast<size_t>(kPayload) * npaths) &&
                 dalloc(st.esc_rec, npaths) && dalloc(st.esc_cz, npaths) && dalloc(st.esc_key, npaths) &&
                 dalloc(st.esc_srec, npaths) && dalloc(st.esc_scz, npaths);
            adj_capacity = npaths;
            release_graph();
        }
        if (ok && sp_capacity > adj_sp_capacity) {
            ok = dalloc(st.sp_x, static_cast<size_t>(kPayload) * sp_capacity);
            adj_sp_capacity = sp_capacity;
            release_graph();
        }
        if (ok && !d_grads) {
            ok = dalloc(d_grads, 1);
            release_graph();
        }
        if (ok && !st.esc_hist) {
            ok = dalloc(st.esc_hist, static_cast<size_t>(kSeg) * kSplatSortParts * kMaxSplatTiles);
            release_graph();
        }
        if (!ok) {
            release();
            (void)hipGetLastError();
            return false;
        }
        st.grads = d_grads;
        return true;
    }

    bool ensure_spill(size_t entries) {
        if (entries <= spill_capacity) return true;
        if (!dalloc(st.spill, entries)) {
            release();
            (void)hipGetLastError();
            return false;
        }
        spill_capacity = entries;
        release_graph();
        return true;
    }

    bool ensure(uint32_t npaths, uint32_t npix, uint32_t nlights, uint32_t nmat, uint32_t njobs) {
        bool ok = true;
        if (npaths > capacity) {
            ok = ok && dalloc(st.path, npaths) && dalloc(st.lo, npaths) && dalloc(st.hit, npaths) &&
                 dalloc(st.q[0], npaths) && dalloc(st.q[1], npaths) && dalloc(st.sorted, npaths) &&
                 dalloc(st.qkey, npaths) &&
                 dalloc(st.sa_o, npaths) && dalloc(st.sa_d, npaths) && dalloc(st.sa_c, npaths) &&
                 dalloc(st.se_o, npaths) && dalloc(st.se_d, npaths) && dalloc(st.se_c, npaths);
            capacity = npaths;
            release_graph();
        }
        const uint64_t spc64 = nlights ? static_cast<uint64_t>(npaths) * nlights : 1u;
        if (spc64 > 0xFFFFFFFFull) {
            release();
            return false;
        }
        const uint32_t spc = static_cast<uint32_t>(spc64);
        if (spc > sp_capacity) {
            ok = ok && dalloc(st.sp_o, spc) && dalloc(st.sp_d, spc) && dalloc(st.sp_c, spc);
            sp_capacity = spc;
            release_graph();
        }
        if (npix > pixels) {
            ok = ok && dalloc(st.sum, npix);
            pixels = npix;
            release_graph();
        }
        const uint32_t nb = std::min<uint32_t>(nmat + 1, kMaxSortBins);  // materials + the miss bin
        if (nb > nbins) {
            ok = ok && dalloc(st.blockhist, static_cast<size_t>(kSeg) * kSortParts * nb);
            nbins = nb;
            release_graph();
        }
        if (njobs > job_capacity) {
            ok = ok && dalloc(d_jobs, njobs);
            job_capacity = njobs;
            release_graph();
        }
        if (!st.ctrl) ok = ok && dalloc(st.ctrl, C_WORDS);
        if (!ok) {
            release();
            (void)hipGetLastError();  // the failed allocation is reported by the caller, not by the next launch
            return false;
        }
        st.jobs = d_jobs;
        st.capacity = capacity;
        st.sp_capacity = sp_capacity;
        return ok;
    }
};

Wavefront* wavefront_create() { return new Wavefront(); }
void wavefront_destroy(Wavefront* w) { delete w; }

namespace {

int grid_for(const void* kernel, int block, int num_cus, int waves_per_simd) {
    int per_cu = 0;
    if (waves_per_simd > 0) {
        per_cu = std::max(1, waves_per_simd * 4 * 64 / block);
    } else if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess || per_cu <= 0) {
        per_cu = 1;
    }
    return num_cus * per_cu;
}

// What a chunk's launches depend on besides device memory.
struct ChunkShape {
    int max_depth;
    uint32_t nbins;
    bool sort, lights;
    bool area, env;  // some job samples emissive triangles / the environment (else that shadow queue stays empty)
    bool adjoint;    // primal pass, then the replay of the path-replay adjoint
    bool kept;       // adjoint: the primal radiance is kept from the forward batch (no primal pass)
    int adj_np;      // adjoint: partials of the BSDF derivatives (3 without roughness/metallic gradients, else 5)
    bool wide;       // traverse the 4-wide BVH (every job's scene has one)
    uint32_t env_tiles;  // adjoint with an environment gradient: splat tiles of the escapes' gradients (else 0)
};

// Enqueues one chunk: generate, max_depth bounces, finalize.  Grid sizes are
// computed beforehand (no runtime queries while a graph is being captured).
struct Grids {
    int extend, shade, escape;
    int ashade, ashadow, aescape, ashade3;  // the adjoint's replay kernels (ashade3: base colour only)
};

// The traversal kernels of one stack variant (LDS = stack entries in LDS)
// over the binary or the 4-wide BVH.
template <bool COUNT, int EB, int LDS, bool WIDE>
void launch_traversal_extend(const Grids& g, hipStream_t stream, const WfState& w, int in, DevCounters* ctr) {
    hipLaunchKernelGGL((k_wf_extend<COUNT, EB, LDS, WIDE>), dim3(g.extend), dim3(EB), 0, stream, w, in, ctr);
}
template <bool COUNT, int EB, int LDS, bool WIDE>
void launch_traversal_shadow(const Grids& g, hipStream_t stream, const WfState& w, const ChunkShape& cs,
                             DevCounters* ctr) {
    if (cs.area)
        hipLaunchKernelGGL((k_wf_shadow<COUNT, EB, LDS, WIDE>), dim3(g.extend), dim3(EB), 0, stream, w, C_SACNT,
                           C_CUR_SA, w.sa_o, w.sa_d, w.sa_c, ctr);
    if (cs.env)
        hipLaunchKernelGGL((k_wf_shadow<COUNT, EB, LDS, WIDE>), dim3(g.extend), dim3(EB), 0, stream, w, C_SECNT,
                           C_CUR_SE, w.se_o, w.se_d, w.se_c, ctr);
    if (cs.lights)
        hipLaunchKernelGGL((k_wf_shadow_punctual<COUNT, EB, LDS, WIDE>), dim3(g.extend), dim3(EB), 0, stream, w,
                           ctr);
}

template <bool COUNT, int EB, bool WIDE>
void launch_replay_shadow(const Grids& g, hipStream_t stream, const WfState& w, const ChunkShape& cs,
                          DevCounters* ctr) {
    if (cs.area)
        hipLaunchKernelGGL((k_wfa_shadow<COUNT, EB, false, WIDE>), dim3(g.ashadow), dim3(EB), 0, stream, w, C_SACNT,
                           C_CUR_SA, w.sa_o, w.sa_d, w.sa_c, w.sa_x, ctr);
    if (cs.env)
        hipLaunchKernelGGL((k_wfa_shadow<COUNT, EB, true, WIDE>), dim3(g.ashadow), dim3(EB), 0, stream, w, C_SECNT,
                           C_CUR_SE, w.se_o, w.se_d, w.se_c, w.se_x, ctr);
    if (cs.lights)
        hipLaunchKernelGGL((k_wfa_shadow_punctual<COUNT, EB, WIDE>), dim3(g.ashadow), dim3(EB), 0, stream, w, ctr);
}

template <bool COUNT, int EB>
void launch_extend(const WavefrontConfig& cfg, const ChunkShape& cs, const Grids& g, hipStream_t stream,
                   const WfState& w, int in, DevCounters* ctr) {
    if (cs.wide) {
        if (cfg.stack_lds == 16) launch_traversal_extend<COUNT, EB, 16, true>(g, stream, w, in, ctr);
        else launch_traversal_extend<COUNT, EB, kStackSize, true>(g, stream, w, in, ctr);
    } else {
        if (cfg.stack_lds == 16) launch_traversal_extend<COUNT, EB, 16, false>(g, stream, w, in, ctr);
        else launch_traversal_extend<COUNT, EB, kStackSize, false>(g, stream, w, in, ctr);
    }
}

template <bool COUNT, int EB>
void launch_shadow(const WavefrontConfig& cfg, const ChunkShape& cs, const Grids& g, hipStream_t stream,
                   const WfState& w, bool replay, DevCounters* ctr) {
    if (replay) {
        if (cs.wide) launch_replay_shadow<COUNT, EB, true>(g, stream, w, cs, ctr);
        else launch_replay_shadow<COUNT, EB, false>(g, stream, w, cs, ctr);
    } else if (cs.wide) {
        if (cfg.stack_lds == 16) launch_traversal_shadow<COUNT, EB, 16, true>(g, stream, w, cs, ctr);
        else launch_traversal_shadow<COUNT, EB, kStackSize, true>(g, stream, w, cs, ctr);
    } else {
        if (cfg.stack_lds == 16) launch_traversal_shadow<COUNT, EB, 16, false>(g, stream, w, cs, ctr);
        else launch_traversal_shadow<COUNT, EB, kStackSize, false>(g, stream, w, cs, ctr);
    }
}

// Camera rays and max_depth bounces of the chunk's paths; `replay` runs the
// adjoint's shading and shadow kernels.
template <bool COUNT, int EB, int SB>
void enqueue_paths(Wavefront& wf, const WavefrontConfig& cfg, const ChunkShape& cs, const Grids& g,
                   DevCounters* ctr, hipStream_t stream, StageTimer* tm, bool replay) {
    const WfState& w = wf.st;
    const size_t sort_lds = sizeof(uint32_t) * cs.nbins;
    {
        Timed t(tm, WF_CONTROL);
        hipLaunchKernelGGL(k_wf_chunk_begin, dim3(1), dim3(64), 0, stream, w);
    }
    {
        Timed t(tm, WF_GENERATE);
        hipLaunchKernelGGL(k_wf_generate, dim3((w.chunk_paths + 255) / 256), dim3(256), 0, stream, w);
    }
    for (int d = 0; d < cs.max_depth; ++d) {
        const int in = d & 1;
        {
            Timed t(tm, WF_CONTROL);
            hipLaunchKernelGGL(k_wf_bounce_begin, dim3(1), dim3(128), 0, stream, w.ctrl, in ^ 1);
        }
        {
            Timed t(tm, WF_EXTEND);
            launch_extend<COUNT, EB>(cfg, cs, g, stream, w, in, ctr);
        }
        if (cs.sort) {
            Timed t(tm, WF_SORT);
            const SortSpec m{w.qkey,      w.blockhist, w.sorted, w.q[in], w.ctrl + C_QCNT + in * kSeg * C_LINE,
                             w.ctrl + C_MISS0, kSortParts};
            hipLaunchKernelGGL(k_wf_sort_count, dim3(kSeg * kSortParts), dim3(256), sort_lds, stream, w, m, cs.nbins);
            if (static_cast<size_t>(cs.nbins) * kSortParts > 65536)  // many materials: more scan threads
                hipLaunchKernelGGL(k_wf_sort_scan<1024>, dim3(kSeg), dim3(1024), 0, stream, w, m, cs.nbins);
            else
                hipLaunchKernelGGL(k_wf_sort_scan<256>, dim3(kSeg), dim3(256), 0, stream, w, m, cs.nbins);
            hipLaunchKernelGGL(k_wf_sort_scatter, dim3(kSeg * kSortParts), dim3(256), sort_lds, stream, w, m,
                               cs.nbins);
        }
        {
            Timed t(tm, WF_SHADE);
            const int us = cs.sort ? 1 : 0;
            const int split = cs.sort ? 1 : 0;  // escapes to k_wf_shade_escape
            if (replay) {
                if (cs.adj_np == 3)
                    hipLaunchKernelGGL((k_wfa_shade<SB, 3>), dim3(g.ashade3), dim3(SB), 0, stream, w, in, us, split);
                else
                    hipLaunchKernelGGL((k_wfa_shade<SB, kDP>), dim3(g.ashade), dim3(SB), 0, stream, w, in, us,
                                       split);
                if (split)
                    hipLaunchKernelGGL((k_wfa_shade_escape<SB>), dim3(g.aescape), dim3(SB), 0, stream, w, in);
            }
            else if (cfg.shade_waves == 4)
                hipLaunchKernelGGL((k_wf_shade<SB, 4>), dim3(g.shade), dim3(SB), 0, stream, w, in, us, split);
            else if (cfg.shade_waves == 3)
                hipLaunchKernelGGL((k_wf_shade<SB, 3>), dim3(g.shade), dim3(SB), 0, stream, w, in, us, split);
            else
                hipLaunchKernelGGL((k_wf_shade<SB, 0>), dim3(g.shade), dim3(SB), 0, stream, w, in, us, split);
            if (!replay && split)
                hipLaunchKernelGGL((k_wf_shade_escape<SB>), dim3(g.escape), dim3(SB), 0, stream, w, in);
        }
        {
            Timed t(tm, WF_SHADOW);
            launch_shadow<COUNT, EB>(cfg, cs, g, stream, w, replay, ctr);
        }
        if (replay && cs.sort && cs.env_tiles) {
            // the bounce's environment-gradient records (escapes after the
            // first bounce, unoccluded environment shadow rays: at most one
            // per path) to the gradient
            Timed t(tm, WF_SHADE);
            const SortSpec e{w.esc_key, w.esc_hist, nullptr, nullptr, w.ctrl + C_ESCNT, nullptr, kSplatSortParts};
            const size_t tile_lds = sizeof(uint32_t) * cs.env_tiles;
            hipLaunchKernelGGL(k_wf_sort_count, dim3(kSeg * kSplatSortParts), dim3(256), tile_lds, stream, w, e,
                               cs.env_tiles);
            hipLaunchKernelGGL(k_wf_sort_scan<1024>, dim3(kSeg), dim3(1024), 0, stream, w, e, cs.env_tiles);
            hipLaunchKernelGGL(k_wfa_env_scatter, dim3(kSeg * kSplatSortParts), dim3(256), tile_lds, stream, w, e,
                               cs.env_tiles);
            hipLaunchKernelGGL(k_wfa_env_splat, dim3(kSeg * cs.env_tiles * kSplatParts), dim3(256), 0, stream, w);
        }
    }
}

// One chunk: its paths (for the adjoint: the primal pass, the seed of the
// replay and the replay), then the per-pixel sums.
template <bool COUNT, int EB, int SB>
void enqueue_chunk(Wavefront& wf, const WavefrontConfig& cfg, const ChunkShape& cs, const Grids& g,
                   DevCounters* ctr, hipStream_t stream, StageTimer* tm, uint32_t gp0, uint32_t gp1) {
    const WfState& w = wf.st;
    if (!(cs.adjoint && cs.kept)) enqueue_paths<COUNT, EB, SB>(wf, cfg, cs, g, ctr, stream, tm, false);
    if (cs.adjoint) {
        {
            Timed t(tm, WF_CONTROL);
            hipLaunchKernelGGL(k_wfa_seed, dim3((w.chunk_paths + 255) / 256), dim3(256), 0, stream, w);
        }
        enqueue_paths<COUNT, EB, SB>(wf, cfg, cs, g, ctr, stream, tm, true);
    }
    {
        Timed t(tm, WF_CONTROL);
        // pixels touched by the chunk: [gp0, gp1) when known (eager), else all
        const uint32_t np = gp1 - gp0;
        hipLaunchKernelGGL(k_wf_finalize, dim3((np + kFinBlock - 1) / kFinBlock), dim3(kFinBlock), 0, stream, w, gp0,
                           gp1);
        hipLaunchKernelGGL(k_wf_chunk_end, dim3(1), dim3(64), 0, stream, w.ctrl);
    }
}

template <bool COUNT, int EB, bool WIDE>
int traversal_grid(const Wavefront& wf, const WavefrontConfig& cfg) {
    // extend and shadow kernels share the grid: take the smaller occupancy
    const void* ext = cfg.stack_lds == 16 ? reinterpret_cast<const void*>(&k_wf_extend<COUNT, EB, 16, WIDE>)
                                          : reinterpret_cast<const void*>(&k_wf_extend<COUNT, EB, kStackSize, WIDE>);
    const void* shd = cfg.stack_lds == 16 ? reinterpret_cast<const void*>(&k_wf_shadow<COUNT, EB, 16, WIDE>)
                                          : reinterpret_cast<const void*>(&k_wf_shadow<COUNT, EB, kStackSize, WIDE>);
    return std::min(grid_for(ext, EB, wf.num_cus, cfg.waves_per_simd),
                    grid_for(shd, EB, wf.num_cus, cfg.waves_per_simd));
}

template <bool COUNT, int EB, int SB>
Grids grids_for(const Wavefront& wf, const WavefrontConfig& cfg, bool wide) {
    Grids g;
    g.extend = wide ? traversal_grid<COUNT, EB, true>(wf, cfg) : traversal_grid<COUNT, EB, false>(wf, cfg);
    const void* shade = cfg.shade_waves == 4   ? reinterpret_cast<const void*>(&k_wf_shade<SB, 4>)
                        : cfg.shade_waves == 3 ? reinterpret_cast<const void*>(&k_wf_shade<SB, 3>)
                                               : reinterpret_cast<const void*>(&k_wf_shade<SB, 0>);
    g.shade = grid_for(shade, SB, wf.num_cus, cfg.waves_per_simd);
    g.escape = grid_for(reinterpret_cast<const void*>(&k_wf_shade_escape<SB>), SB, wf.num_cus, cfg.waves_per_simd);
    g.ashade = grid_for(reinterpret_cast<const void*>(&k_wfa_shade<SB, kDP>), SB, wf.num_cus, cfg.waves_per_simd);
    g.ashade3 = grid_for(reinterpret_cast<const void*>(&k_wfa_shade<SB, 3>), SB, wf.num_cus, cfg.waves_per_simd);
    g.aescape =
        grid_for(reinterpret_cast<const void*>(&k_wfa_shade_escape<SB>), SB, wf.num_cus, cfg.waves_per_simd);
    g.ashadow = wide ? std::min(grid_for(reinterpret_cast<const void*>(&k_wfa_shadow<COUNT, EB, true, true>), EB,
                                         wf.num_cus, cfg.waves_per_simd),
                                grid_for(reinterpret_cast<const void*>(&k_wfa_shadow_punctual<COUNT, EB, true>), EB,
                                         wf.num_cus, cfg.waves_per_simd))
                     : std::min(grid_for(reinterpret_cast<const void*>(&k_wfa_shadow<COUNT, EB, true, false>), EB,
                                         wf.num_cus, cfg.waves_per_simd),
                                grid_for(reinterpret_cast<const void*>(&k_wfa_shadow_punctual<COUNT, EB, false>), EB,
                                         wf.num_cus, cfg.waves_per_simd));
    return g;
}

// Calls f.template operator()<EB, SB>() for the configured block sizes.
template <class F>
void with_blocks(const WavefrontConfig& cfg, F&& f) {
    const int eb = cfg.extend_block, sb = cfg.shade_block;
    if (eb == 64 && sb == 64) f.template operator()<64, 64>();
    else if (eb == 64 && sb == 128) f.template operator()<64, 128>();
    else if (eb == 64 && sb == 256) f.template operator()<64, 256>();
    else if (eb == 128 && sb == 64) f.template operator()<128, 64>();
    else if (eb == 128 && sb == 256) f.template operator()<128, 256>();
    else if (eb == 256 && sb == 64) f.template operator()<256, 64>();
    else if (eb == 256 && sb == 128) f.template operator()<256, 128>();
    else if (eb == 256 && sb == 256) f.template operator()<256, 256>();
    else f.template operator()<128, 128>();
}

// Global pixels touched by global paths [c0, c1) (host copy of the job table).
void chunk_pixels(const std::vector<JobDesc>& jobs, unsigned long long c0, unsigned long long c1, uint32_t& gp0,
                  uint32_t& gp1) {
    auto pixel_of = [&](unsigned long long g) -> uint32_t {
        size_t j = 0;
        while (j + 1 < jobs.size() && jobs[j + 1].path_offset <= g) ++j;
        const JobDesc& d = jobs[j];
        const unsigned long long px = (g - d.path_offset) / static_cast<unsigned long long>(d.params.spf);
        return d.pixel_offset + static_cast<uint32_t>(std::min<unsigned long long>(px, d.npix));
    };
    gp0 = pixel_of(c0);
    gp1 = c1 > c0 ? pixel_of(c1 - 1) + 1 : gp0;
}

template <bool COUNT>
struct ChunkEnqueuer {
    Wavefront& wf;
    const WavefrontConfig& cfg;
    const ChunkShape& cs;
    const std::vector<JobDesc>& host_jobs;
    DevCounters* ctr;
    hipStream_t stream;
    StageTimer* tm;
    int nchunks;
    const Grids& g;
    template <int EB, int SB>
    void operator()() {
        for (int c = 0; c < nchunks; ++c) {
            const unsigned long long c0 = static_cast<unsigned long long>(c) * wf.st.chunk_paths;
            const unsigned long long c1 = std::min(c0 + wf.st.chunk_paths, wf.st.total_paths);
            uint32_t gp0, gp1;
            chunk_pixels(host_jobs, c0, c1, gp0, gp1);
            enqueue_chunk<COUNT, EB, SB>(wf, cfg, cs, g, ctr, stream, tm, gp0, gp1);
        }
    }
};

template <bool COUNT>
struct GridQuery {
    const Wavefront& wf;
    const WavefrontConfig& cfg;
    bool wide;
    Grids out;
    template <int EB, int SB>
    void operator()() {
        out = grids_for<COUNT, EB, SB>(wf, cfg, wide);
    }
};

// The captured chunk covers every pixel (the finalize kernel skips pixels the
// chunk does not touch), so one graph serves every chunk.
template <bool COUNT>
struct CaptureEnqueuer {
    Wavefront& wf;
    const WavefrontConfig& cfg;
    const ChunkShape& cs;
    const Grids& g;
    DevCounters* ctr;
    hipStream_t stream;
    template <int EB, int SB>
    void operator()() {
        enqueue_chunk<COUNT, EB, SB>(wf, cfg, cs, g, ctr, stream, nullptr, 0u, wf.st.total_pixels);
    }
};

}  // namespace

namespace {

// One batch of every job: the forward render (G == null) or the adjoint.
// What a kept radiance array was rendered from: the caller's key (which
// changes with every update of the scene's device data), then each job's
// scene view and batch parameters (camera, frame, samples, depth, sampler).
std::vector<unsigned char> radiance_signature(uint64_t key, const std::vector<WavefrontJob>& jobs) {
    std::vector<unsigned char> sig(sizeof(key) + jobs.size() * (sizeof(DevScene) + sizeof(RenderParams)));
    unsigned char* o = sig.data();
    std::memcpy(o, &key, sizeof(key));
    o += sizeof(key);
    for (const WavefrontJob& j : jobs) {
        std::memcpy(o, &j.scene, sizeof(DevScene));
        o += sizeof(DevScene);
        std::memcpy(o, &j.params, sizeof(RenderParams));
        o += sizeof(RenderParams);
    }
    return sig;
}

hipError_t run_batch(Wavefront* wf, const std::vector<WavefrontJob>& jobs_in, const WavefrontConfig& cfg,
                     DevCounters* ctr, bool count, hipStream_t stream, float* ms, std::string* err,
                     const AdjointGrads* G, uint64_t radiance_key) {
    if (jobs_in.empty()) return hipSuccess;
    // job table: global path and pixel offsets
    std::vector<JobDesc> jobs(jobs_in.size());
    unsigned long long paths = 0, pixels = 0;
    uint32_t max_lights = 0, max_mat = 0;
    int max_depth = 0;
    bool any_area = false, any_env = false, all_wide = true;
    for (size_t j = 0; j < jobs_in.size(); ++j) {
        const WavefrontJob& in = jobs_in[j];
        JobDesc& d = jobs[j];
        std::memset(&d, 0, sizeof(d));
        d.scene = in.scene;
        d.params = in.params;
        d.film_new = in.film_new;
        d.film_acc = in.film_acc;
        d.img_grad = in.img_grad;
        d.replay = in.replay;
        if (G && (in.img_grad == nullptr || std::memcmp(&in.scene, &jobs_in[0].scene, sizeof(DevScene)) != 0)) {
            if (err) *err = "adjoint jobs need an image gradient each and must be views of one scene";
            return hipErrorInvalidValue;
        }
        const unsigned long long npix = static_cast<unsigned long long>(in.params.width) * in.params.height;
        if (in.params.spf <= 0 || npix == 0) {
            if (err) *err = "wavefront job with no pixels or samples";
            return hipErrorInvalidValue;
        }
        d.path_offset = paths;
        d.pixel_offset = static_cast<uint32_t>(pixels);
        d.npix = static_cast<uint32_t>(npix);
        paths += npix * static_cast<unsigned long long>(in.params.spf);
        pixels += npix;
        max_lights = std::max(max_lights, in.scene.num_lights);
        // the tests bounce() makes before sampling each light type
        any_area = any_area || (in.scene.num_emissive > 0 && in.scene.emissive != nullptr);
        any_env = any_env || (in.scene.env != nullptr && in.scene.env_alias != nullptr && in.scene.num_env_texels > 0 &&
                              in.scene.env_w > 0 && in.scene.env_h > 0);
        max_mat = std::max(max_mat, in.scene.num_materials);
        max_depth = std::max(max_depth, in.params.max_depth);
        all_wide = all_wide && in.scene.wnodes != nullptr && in.scene.num_wnodes > 0;
    }
    if (pixels > 0xFFFFFFFFull) {
        if (err) *err = "wavefront batch has more than 2^32 pixels";
        return hipErrorInvalidValue;
    }
    const unsigned long long want = std::max(1ull, static_cast<unsigned long long>(std::max(1, cfg.chunk_spp)) * pixels);
    if (wf->num_cus == 0) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&wf->num_cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            wf->num_cus <= 0)
            wf->num_cus = 256;
    }
    const uint32_t njobs = static_cast<uint32_t>(jobs.size());
    // At most 2^28 paths in flight (~56 GB of path state and queues, 208 B
    // per path; the adjoint's replay state, pending records, shadow payloads
    // and splat records ~68 GB more, 252 B per path): a 1080p frame of 64
    // samples runs as one chunk, whose bounce loop drains once (every chunk
    // ends in a tail of nearly empty launches).  A device that cannot hold
    // that many gets half as many, down to 2^22.
    unsigned long long cap = std::min<unsigned long long>(1ull << 28, want);
    uint32_t C = 0;
    while (true) {
        C = static_cast<uint32_t>(std::min(cap, paths));
        if (wf->ensure(C, static_cast<uint32_t>(pixels), max_lights, max_mat, njobs) &&
            (!G || wf->ensure_adjoint(C)))
            break;
        if (cap <= (1ull << 22)) {
            if (err) *err = "wavefront: out of device memory for path state";
            return hipErrorOutOfMemory;
        }
        cap /= 2;
    }
    const int nchunks = static_cast<int>((paths + C - 1) / C);
    wf->st.njobs = njobs;
    wf->st.chunk_paths = C;
    wf->st.grab_extend = static_cast<uint32_t>(cfg.grab_extend);
    wf->st.grab_shade = static_cast<uint32_t>(cfg.grab_shade);
    wf->st.grab_shadow = static_cast<uint32_t>(cfg.grab_shadow);
    wf->st.shade_ctr = count ? ctr : nullptr;
    wf->st.total_paths = paths;
    wf->st.total_pixels = static_cast<uint32_t>(pixels);
    ChunkShape cs;
    cs.max_depth = max_depth;
    cs.nbins = max_mat + 1;
    cs.sort = cfg.sort_by_material && cs.nbins <= kMaxSortBins;
    cs.lights = max_lights > 0;
    cs.area = any_area;
    cs.env = any_env;
    cs.adjoint = G != nullptr;
    cs.adj_np = G && !G->rough && !G->metal ? 3 : kDP;
    cs.wide = cfg.bvh_width == 4 && all_wide;
    {
        // the adjoint's jobs are views of one scene: its environment's tiles
        const DevScene& s0 = jobs_in[0].scene;
        const uint32_t tx = s0.env_w > 0 ? (static_cast<uint32_t>(s0.env_w) + kSplatTileW - 1) / kSplatTileW : 0u;
        const uint32_t ty = s0.env_h > 0 ? (static_cast<uint32_t>(s0.env_h) + kSplatTileH - 1) / kSplatTileH : 0u;
        const bool splat = G != nullptr && G->env != nullptr && s0.env != nullptr && cs.sort && tx * ty > 0 &&
                           tx * ty <= kMaxSplatTiles;
        cs.env_tiles = splat ? tx * ty : 0u;
        wf->st.env_tiles_x = tx;
        wf->st.env_tiles = cs.env_tiles;
    }
    wf->st.key_bins = cs.sort ? cs.nbins : 0u;
    // kept radiance: a forward with a key writes each path's radiance (16 B
    // per path of the batch; without the memory for it the batch renders
    // without keeping), an adjoint whose batch signature matches reads it
    wf->st.kept_mode = 0;
    wf->st.kept = nullptr;
    wf->last_reused = false;
    std::vector<unsigned char> sig;
    if (radiance_key != 0) sig = radiance_signature(radiance_key, jobs_in);
    if (radiance_key != 0 && !G) {
        wf->kept_sig.clear();  // overwritten below (or not kept at all)
        if (paths > wf->kept_capacity) {
            wf->release_kept();
            if (dalloc(wf->kept, paths)) {
                wf->kept_capacity = paths;
            } else {
                wf->kept = nullptr;
                (void)hipGetLastError();
            }
        }
        if (wf->kept) wf->st.kept_mode = 1;
    } else if (radiance_key != 0 && G && wf->kept && !wf->kept_sig.empty() && wf->kept_sig == sig) {
        wf->st.kept_mode = 2;
        wf->last_reused = true;
    }
    wf->st.kept = wf->st.kept_mode ? wf->kept : nullptr;
    cs.kept = wf->st.kept_mode == 2;
    hipError_t e;
    if ((e = hipMemcpyAsync(wf->d_jobs, jobs.data(), sizeof(JobDesc) * njobs, hipMemcpyHostToDevice, stream)) !=
        hipSuccess)
        return e;
    if (G && (e = hipMemcpyAsync(wf->d_grads, G, sizeof(AdjointGrads), hipMemcpyHostToDevice, stream)) != hipSuccess)
        return e;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, stream);
    hipLaunchKernelGGL(k_wf_batch_begin, dim3(1), dim3(64), 0, stream, wf->st.ctrl);
    const bool use_graph = cfg.use_graph && !cfg.time_stages;
    uint64_t key[14] = {static_cast<uint64_t>(C), static_cast<uint64_t>(paths), static_cast<uint64_t>(pixels),
                        static_cast<uint64_t>(njobs), static_cast<uint64_t>(max_depth),
                        static_cast<uint64_t>(cfg.extend_block * 1000 + cfg.shade_block),
                        static_cast<uint64_t>(cs.sort * 100 + cfg.waves_per_simd + 1000 * cfg.shade_waves +
                                              100000 * cs.adj_np) |
                            (static_cast<uint64_t>(cs.env_tiles) << 32) | (static_cast<uint64_t>(wf->st.env_tiles_x) << 48),
                        static_cast<uint64_t>(count ? 1 : 0), reinterpret_cast<uint64_t>(ctr),
                        static_cast<uint64_t>(cs.lights + 2 * cs.area + 4 * cs.env + 8 * cs.wide),
                        static_cast<uint64_t>(cs.nbins),
                        static_cast<uint64_t>(cfg.stack_lds + (cs.adjoint ? 1000 : 0)) |
                            (static_cast<uint64_t>(cfg.grab_extend) << 16) | (static_cast<uint64_t>(cfg.grab_shade) << 32) |
                            (static_cast<uint64_t>(cfg.grab_shadow) << 48),
                        reinterpret_cast<uint64_t>(wf->st.kept), static_cast<uint64_t>(wf->st.kept_mode)};
    StageTimer timer;
    timer.stream = stream;
    StageTimer* tm = cfg.time_stages ? &timer : nullptr;
    Grids g;
    if (count) {
        GridQuery<true> q{*wf, cfg, cs.wide, {}};
        with_blocks(cfg, q);
        g = q.out;
    } else {
        GridQuery<false> q{*wf, cfg, cs.wide, {}};
        with_blocks(cfg, q);
        g = q.out;
    }
    // traversal stack entries kept in global memory: one column per thread of the traversal grid
    // (the replay's shadow kernels keep kReplayStackLds entries in LDS)
    const int total = cs.wide ? kWideStack : kStackSize;
    const size_t spill = std::max(static_cast<size_t>(total - cfg.stack_lds) * g.extend * cfg.extend_block,
                                  cs.adjoint ? static_cast<size_t>(total - kReplayStackLds) * g.ashadow *
                                                   cfg.extend_block
                                             : size_t(0));
    if (spill > 0 && !wf->ensure_spill(spill)) {
        if (err) *err = "wavefront: out of device memory for the traversal stacks";
        return hipErrorOutOfMemory;
    }
    if (use_graph) {
        const int gk = cs.adjoint ? 1 : 0;
        if (!wf->exec[gk] || std::memcmp(key, wf->key[gk], sizeof(key)) != 0) {
            wf->release_graph(gk);
            hipStream_t cap_stream;
            if ((e = hipStreamCreate(&cap_stream)) != hipSuccess) return e;
            e = hipStreamBeginCapture(cap_stream, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess) {
                if (count) {
                    CaptureEnqueuer<true> f{*wf, cfg, cs, g, ctr, cap_stream};
                    with_blocks(cfg, f);
                } else {
                    CaptureEnqueuer<false> f{*wf, cfg, cs, g, ctr, cap_stream};
                    with_blocks(cfg, f);
                }
                e = hipStreamEndCapture(cap_stream, &wf->graph[gk]);
            }
            if (e == hipSuccess) e = hipGraphInstantiate(&wf->exec[gk], wf->graph[gk], nullptr, nullptr, 0);
            (void)hipStreamDestroy(cap_stream);
            if (e != hipSuccess) {
                wf->release_graph(gk);
                if (err) *err = std::string("wavefront graph capture: ") + hipGetErrorString(e);
                return e;
            }
            std::memcpy(wf->key[gk], key, sizeof(key));
        }
        for (int c = 0; c < nchunks; ++c)
            if ((e = hipGraphLaunch(wf->exec[gk], stream)) != hipSuccess) return e;
    } else if (count) {
        ChunkEnqueuer<true> f{*wf, cfg, cs, jobs, ctr, stream, tm, nchunks, g};
        with_blocks(cfg, f);
    } else {
        ChunkEnqueuer<false> f{*wf, cfg, cs, jobs, ctr, stream, tm, nchunks, g};
        with_blocks(cfg, f);
    }
    {
        Timed t(tm, WF_CONTROL);
        const dim3 grid((static_cast<uint32_t>(pixels) + 255) / 256);
        if (cs.adjoint) hipLaunchKernelGGL(k_wfa_replay_out, grid, dim3(256), 0, stream, wf->st);
        else hipLaunchKernelGGL(k_wf_film, grid, dim3(256), 0, stream, wf->st, ctr, count ? 1u : 0u);
    }
    (void)hipEventRecord(e1, stream);
    e = hipEventSynchronize(e1);
    float t = 0.0f;
    (void)hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms) *ms = t;
    if (tm) {
        timer.collect(wf->stage_ms, wf->stage_launches);
    } else {
        for (int k = 0; k < WF_STAGES; ++k) {
            wf->stage_ms[k] = 0.0f;
            wf->stage_launches[k] = 0;
        }
    }
    if (e != hipSuccess) return e;
    e = hipGetLastError();
    if (e == hipSuccess && wf->st.kept_mode == 1) wf->kept_sig = std::move(sig);
    return e;
}

}  // namespace

hipError_t wavefront_render(Wavefront* wf, const std::vector<WavefrontJob>& jobs, const WavefrontConfig& cfg,
                            DevCounters* ctr, bool count, hipStream_t stream, float* ms, std::string* err,
                            uint64_t radiance_key) {
    return run_batch(wf, jobs, cfg, ctr, count, stream, ms, err, nullptr, radiance_key);
}

hipError_t wavefront_adjoint(Wavefront* wf, const std::vector<WavefrontJob>& jobs, const AdjointGrads& g,
                             const WavefrontConfig& cfg, DevCounters* ctr, bool count, hipStream_t stream, float* ms,
                             std::string* err, uint64_t radiance_key) {
    return run_batch(wf, jobs, cfg, ctr, count, stream, ms, err, &g, radiance_key);
}

void wavefront_release_kept(Wavefront* w) {
    if (w) w->release_kept();
}

bool wavefront_adjoint_reused_radiance(const Wavefront* w) { return w && w->last_reused; }

void wavefront_stage_times(const Wavefront* w, float* ms, uint32_t* launches) {
    for (int k = 0; k < WF_STAGES; ++k) {
        if (ms) ms[k] = w ? w->stage_ms[k] : 0.0f;
        if (launches) launches[k] = w ? w->stage_launches[k] : 0u;
    }
}

}  // namespace ptd

// Wavefront form of the integrator (host interface).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "adjoint.h"
#include "dev_scene.h"

namespace ptd {

struct WavefrontConfig {
    int chunk_spp = 64;         // paths in flight = chunk_spp x (pixels of all jobs), at most 2^28
    int extend_block = 64;      // threads per block of the traversal kernels
    int shade_block = 256;      // threads per block of the shading kernel
    int waves_per_simd = 0;     // persistent grid size: 0 = from the occupancy calculator
    int sort_by_material = 1;   // counting-sort the shading queue by material
    int use_graph = 1;          // capture the bounce loop of a chunk in a hipGraph
    int time_stages = 0;        // bracket every launch with HIP events (no graph); see wavefront_stage_times
    int shade_waves = 3;        // shading kernel built for at least this many waves per SIMD (0, 3 or 4)
    int stack_lds = 16;         // traversal stack entries in LDS (16 or 32); the rest live in device memory
    // queue items a wave takes per atomic (64 .. 1024, a power of two) in the
    // closest-hit, shading and shadow kernels: larger grabs cost fewer
    // same-address atomics, smaller ones keep the rays in flight coherent and
    // the waves' shares even (with each cursor on a line of its own, the
    // smallest grabs were fastest on the 1080p icosphere frame)
    int grab_extend = 64;
    int grab_shade = 128;
    int grab_shadow = 64;
    // 4: traverse the 4-wide BVH with quantized child boxes when every job's
    // scene has one; 2: the binary BVH (both enter children nearest first;
    // they can resolve an exact-distance tie differently)
    int bvh_width = 4;
};

// Stages timed when WavefrontConfig::time_stages is set.
enum WavefrontStage {
    WF_GENERATE = 0,  // camera rays
    WF_EXTEND = 1,    // closest-hit traversal of the ray queue
    WF_SORT = 2,      // material histogram, scan and scatter
    WF_SHADE = 3,     // surface interaction, BSDF, NEE sampling, next rays
    WF_SHADOW = 4,    // any-hit traversal of the shadow queues
    WF_CONTROL = 5,   // queue resets, chunk bookkeeping, finalize, film
    WF_STAGES = 6
};

// One render job: a scene, the parameters of its batch and its films.
struct WavefrontJob {
    DevScene scene;
    RenderParams params;  // params.spf samples per pixel of frame params.frame_count
    float4* film_new;
    float4* film_acc;
    const float4* img_grad = nullptr;  // adjoint only: d loss / d batch mean image (W*H float4)
    float4* replay = nullptr;          // adjoint only: receives the replay pass's batch mean, or null
};

class Wavefront;

// Owns path-state buffers, queues and the captured graph; rebuilt when sizes
// or settings change.
Wavefront* wavefront_create();
void wavefront_destroy(Wavefront* w);
// Renders one batch of every job (all jobs' paths share the queues of one
// persistent schedule) into their films exactly as the megakernel does
// (batch mean + running mean).  ms = GPU time from HIP events on `stream`.
// radiance_key != 0: also keep each path's radiance sum (16 B per path) for
// an adjoint of the same batch; the key must change whenever the scene's
// device data does.
hipError_t wavefront_render(Wavefront* w, const std::vector<WavefrontJob>& jobs, const WavefrontConfig& cfg,
                            DevCounters* ctr, bool count, hipStream_t stream, float* ms, std::string* err,
                            uint64_t radiance_key = 0);
// Path-replay adjoint of one batch of every job (views of one scene, each
// with its img_grad) in the persistent schedule: per chunk the primal pass
// (each path's radiance), then the replay from the same sampler state, which
// accumulates d loss / d parameter into g.  Films are not touched.  When the
// last forward kept its radiance under the same radiance_key (!= 0) for the
// same jobs (scene views and batch parameters), the replay seeds from it and
// the primal passes are not run.
hipError_t wavefront_adjoint(Wavefront* w, const std::vector<WavefrontJob>& jobs, const AdjointGrads& g,
                             const WavefrontConfig& cfg, DevCounters* ctr, bool count, hipStream_t stream, float* ms,
                             std::string* err, uint64_t radiance_key = 0);
// Frees the kept radiance.
void wavefront_release_kept(Wavefront* w);
// Whether the last adjoint seeded from kept radiance.
bool wavefront_adjoint_reused_radiance(const Wavefront* w);
// GPU milliseconds and launch counts per stage of the last batch rendered
// with time_stages set (zeros otherwise).
void wavefront_stage_times(const Wavefront* w, float* ms, uint32_t* launches);

}  // namespace ptd

// Path-replay adjoint of the integrator (host interface).
#pragma once

#include <hip/hip_runtime.h>

#include "dev_scene.h"

namespace ptd {

// Gradient buffers (device, accumulated with +=, any may be null).
struct AdjointGrads {
    float* base;   // num_materials x 3: the material's baseColor factor
    float* rough;  // num_materials
    float* metal;  // num_materials
    float* tex;    // texel pool x 4 (base-colour textures)
    float* env;    // env_w x env_h x 4
};

// Replays batch `p.frame_count` (p.spf samples per pixel, the forward's
// random numbers) and accumulates d loss / d parameter, given img_grad =
// d loss / d (batch mean image), W*H float4.  replay (optional, W*H float4)
// receives the batch mean re-computed by the replay pass.
hipError_t launch_adjoint(const DevScene& s, const RenderParams& p, const float4* img_grad, const AdjointGrads& g,
                          float4* replay, DevCounters* ctr, bool count, hipStream_t stream);

}  // namespace ptd
